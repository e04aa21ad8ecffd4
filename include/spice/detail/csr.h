// Host-side CSR view (API of reference detail/csr.h): offsets[src+1], targets and an
// optional per-synapse payload, iterated as (target, payload*) pairs.  The engine
// keeps its CSRs on the device; this class serves host code and tests.
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <type_traits>
#include <utility>
#include <vector>

#include "spice/topology.h"
#include "spice/util/assert.h"
#include "spice/util/range.h"
#include "spice/util/stdint.h"
#include "spice/util/type_traits.h"

namespace spice::detail {
template <class T = void>
class csr {
public:
	template <bool Const>
	class iterator_t {
	public:
		using edge_t            = std::conditional_t<Const, T const, T>;
		using iterator_category = std::input_iterator_tag;
		using difference_type   = std::ptrdiff_t;
		using value_type        = std::pair<Int32, edge_t*>;

		constexpr iterator_t() = default;
		constexpr iterator_t(iterator_t<false> const& other) : _dst(other._dst), _edge(other._edge) {}

		constexpr value_type operator*() const {
			if constexpr (std::is_void_v<T>)
				return {*_dst, nullptr};
			else
				return {*_dst, _edge};
		}
		constexpr bool operator==(iterator_t const& o) const { return _dst == o._dst; }
		constexpr bool operator!=(iterator_t const& o) const { return _dst != o._dst; }
		constexpr iterator_t& operator++() {
			_dst++;
			if constexpr (!std::is_void_v<T>) _edge++;
			return *this;
		}
		constexpr iterator_t operator++(int) {
			iterator_t r = *this;
			++*this;
			return r;
		}

	private:
		friend class csr;
		template <bool>
		friend class iterator_t;
		Int32 const* _dst = nullptr;
		[[no_unique_address]] util::optional_t<edge_t*, !std::is_void_v<T>> _edge{};

		constexpr explicit iterator_t(Int32 const* dst) : _dst(dst) {}
		template <class U = T, class = std::enable_if_t<!std::is_void_v<U>>>
		constexpr iterator_t(Int32 const* dst, edge_t* edge) : _dst(dst), _edge(edge) {}
	};
	using iterator       = iterator_t<false>;
	using const_iterator = iterator_t<true>;

	csr(Topology& c, util::seed_seq const& seed) {
		_offsets.resize(c.src_count > 0 ? size_t(c.src_count + 1) : 0);
		_neighbors.resize(size_t(c.size()));
		if constexpr (!std::is_void_v<T>) _edges.resize(_neighbors.size());
		c.generate(_offsets, _neighbors, seed);
		SPICE_INV(std::is_sorted(_offsets.begin(), _offsets.end()));
	}

	util::range_t<iterator> neighbors(Int const src) {
		SPICE_INV(0 <= src && src + 1 < Int(_offsets.size()));
		Int const a = _offsets[src], b = _offsets[src + 1];
		if constexpr (std::is_void_v<T>)
			return {iterator(_neighbors.data() + a), iterator(_neighbors.data() + b)};
		else
			return {iterator(_neighbors.data() + a, _edges.data() + a),
			        iterator(_neighbors.data() + b, _edges.data() + b)};
	}
	util::range_t<const_iterator> neighbors(Int const src) const {
		auto r = const_cast<csr*>(this)->neighbors(src);
		return {r.first, r.last};
	}

	std::vector<Int> const& offsets() const { return _offsets; }
	std::vector<Int32> const& targets() const { return _neighbors; }

private:
	std::vector<Int> _offsets;
	std::vector<Int32> _neighbors;
	[[no_unique_address]] util::optional_t<std::vector<T>, !std::is_void_v<T>> _edges;
};
}  // namespace spice::detail
