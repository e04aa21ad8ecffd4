// Connectivity descriptions (API of reference topology.h / topology.cpp).
//
// fixed_probability is generated on the device by the runtime (snn::connect
// hands its probability and seed to libspice_hip); its generate() over host
// buffers also runs the device generator and copies the result back.  adj_list
// and user-defined topologies generate on the host through an edge_stream and
// are uploaded.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "spice/util/assert.h"
#include "spice/util/random.h"
#include "spice/util/stdint.h"
#include "spice_hip.h"

namespace spice {
// Appends (src, dst) edges in ascending src order into CSR arrays.
class edge_stream {
public:
	edge_stream(std::span<Int> offsets, std::span<Int32> neighbors) :
	_offsets(offsets), _neighbors(neighbors) {}

	edge_stream& operator<<(std::pair<Int32, Int32> const edge) {
		SPICE_PRE(_src < Int(_offsets.size()));
		SPICE_PRE(_dst < Int(_neighbors.size()));
		SPICE_PRE(edge.first < Int(_offsets.size()));
		for (; _src <= edge.first; _src++) _offsets[_src] = _dst;
		_neighbors[_dst++] = edge.second;
		return *this;
	}

	// Closes the row of the last source seen (later rows are not touched).
	void flush() {
		SPICE_INV(_src < Int(_offsets.size()));
		_offsets[_src] = _dst;
		_src           = 0;
		_dst           = 0;
	}

private:
	std::span<Int> _offsets;
	std::span<Int32> _neighbors;
	Int _src = 0;
	Int _dst = 0;
};

struct Topology {
	Int src_count = 0;
	Int dst_count = 0;

	virtual ~Topology() = default;

	Topology& operator()(Int const src_count_, Int const dst_count_) {
		SPICE_INV(0 <= src_count_ && src_count_ < std::numeric_limits<Int32>::max());
		SPICE_INV(0 <= dst_count_ && dst_count_ < std::numeric_limits<Int32>::max());
		src_count = src_count_;
		dst_count = dst_count_;
		return *this;
	}

	virtual Int size() const = 0;

	virtual void generate(edge_stream&, util::seed_seq const&) {
		SPICE_PRE(false && "Topology subclasses must implement generate(edge_stream, seed_seq)");
	}

	virtual void generate(std::span<Int> offsets, std::span<Int32> neighbors,
	                      util::seed_seq const& seed) {
		SPICE_PRE(Int(offsets.size()) > src_count);
		SPICE_PRE(Int(neighbors.size()) >= size());
		edge_stream es(offsets, neighbors);
		generate(es, seed);
		es.flush();
	}
};

class adj_list : public Topology {
public:
	void connect(Int const src, Int const dst) {
		SPICE_PRE(0 <= src && src < std::numeric_limits<Int32>::max());
		SPICE_PRE(0 <= dst && dst < std::numeric_limits<Int32>::max());
		_edges.push_back(UInt(src) << 32 | UInt(dst));
	}

	Int size() const override { return Int(_edges.size()); }

	void generate(edge_stream& stream, util::seed_seq const&) override {
		std::sort(_edges.begin(), _edges.end());
		for (UInt const e : _edges) stream << std::pair<Int32, Int32>{Int32(e >> 32), Int32(e & 0xffffffffu)};
	}

private:
	std::vector<UInt> _edges;
};

namespace detail {
inline void check_spice(int status) {
	if (status == SPICE_OK) return;
	if (status == SPICE_EPRECONDITION) throw std::logic_error(spice_last_error());
	throw std::runtime_error(std::string("spice: ") + spice_last_error());
}
}  // namespace detail

class fixed_probability : public Topology {
public:
	explicit fixed_probability(double const p) : _p(p) { SPICE_PRE(0 <= p && p <= 1); }

	double probability() const { return _p; }

	// src_count * max_degree, max_degree = dst*p + 3 sqrt(dst*p*(1-p)) truncated.
	Int size() const override {
		Int const max_degree =
		    Int(double(dst_count) * _p + 3 * std::sqrt(double(dst_count) * _p * (1 - _p)));
		return src_count * max_degree;
	}

	// Runs the device generator on a temporary context and copies the CSR back.
	void generate(std::span<Int> offsets, std::span<Int32> neighbors,
	              util::seed_seq const& seed) override {
		SPICE_PRE(Int(offsets.size()) > src_count);
		SPICE_PRE(Int(neighbors.size()) >= size());
		if (src_count == 0 || dst_count == 0 || _p == 0) return;
		spice_ctx* ctx = nullptr;
		detail::check_spice(spice_ctx_create(0, 1, &ctx));
		struct guard {
			spice_ctx* c;
			~guard() { spice_ctx_destroy(c); }
		} g{ctx};
		int64_t src = 0, dst = 0, proj = 0;
		detail::check_spice(spice_pop_create(ctx, src_count, 0, 0, &src));
		detail::check_spice(spice_pop_create(ctx, dst_count, 1, 0, &dst));
		UInt128 const s = seed.seed();
		detail::check_spice(spice_proj_create_fixed_probability(ctx, src, dst, _p, s.lo, s.hi, 1, &proj));
		int64_t nnz = 0;
		detail::check_spice(spice_proj_synapses(ctx, proj, &nnz));
		std::vector<int64_t> off(size_t(src_count + 1));
		std::vector<int32_t> nb(size_t(std::max<int64_t>(nnz, 1)));
		detail::check_spice(spice_proj_download(ctx, proj, off.data(), nb.data()));
		std::copy(off.begin(), off.end(), offsets.begin());
		std::copy(nb.begin(), nb.begin() + nnz, neighbors.begin());
	}

private:
	double const _p;
};
}  // namespace spice
