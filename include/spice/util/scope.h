// Runs a callable at scope exit (API of reference util/scope.h).
#pragma once

#include <functional>
#include <utility>

namespace spice::util {
class scope_exit {
public:
	scope_exit(std::function<void()> fn) : _fn(std::move(fn)) {}
	~scope_exit() {
		if (_fn) _fn();
	}
	scope_exit(scope_exit const&)            = delete;
	scope_exit& operator=(scope_exit const&) = delete;

private:
	std::function<void()> _fn;
};
}  // namespace spice::util
