// Serving: a request loop that runs simulation jobs of mixed models and sizes
// on one GPU with continuous batching (an extension; the reference runs one
// network per process).
//
// Jobs arrive over time.  A job is admitted as soon as it has arrived and the
// device memory it needs fits in the budget; it then becomes one instance of a
// shared context (spice::batch) and steps together with every other admitted job
// of its class, its state staying in device memory between windows.  A job that
// has run its steps hands its network to the job's `finish` callback and is
// freed, making room for the next.  Two classes run side by side on their own
// streams: networks whose projections all have long delays (multi-step blocks)
// and networks that need one-step blocks (delay-1 projections or per-synapse
// state).  Each job's results are exactly those of the same network run alone:
// instances never interact and every kernel is per instance.
#pragma once

#include <algorithm>
#include <chrono>
#include <functional>
#include <memory>
#include <thread>
#include <utility>
#include <vector>

#include "spice/snn.h"

namespace spice {

struct request {
	float dt        = 1e-4f;
	float max_delay = 15e-4f;
	util::seed_seq seed{{1337}};
	Int steps = 0;
	int64_t footprint = 0;     // device bytes the network holds (admission by memory)
	bool single_step  = false; // needs one-step blocks (delay-1 projections or per-synapse state)
	double arrival    = 0;     // seconds after run() starts
	std::function<void(snn&)> build;
	std::function<void(snn&)> finish;
};

class server {
public:
	struct record {
		double arrival, admitted, finished;  // seconds after run() starts
		Int steps;
	};

	// budget: device bytes the admitted jobs may hold at once; window: steps per
	// round (a job's last round may be shorter); max_delay_steps bounds every
	// job's max_delay / dt.
	// graphs: rounds between admissions and retirements replay the window's
	// launches from HIP graphs (off: every launch is issued, e.g. for tracing).
	server(int64_t budget, Int window = 100, Int max_delay_steps = 64, bool graphs = true) :
	_budget(budget), _window(window),
	_multi(std::make_unique<batch>(max_delay_steps)), _single(std::make_unique<batch>(max_delay_steps)) {
		_multi->set_window(window);
		_single->set_window(window);
		_multi->set_graphs(graphs);
		_single->set_graphs(graphs);
	}

	void submit(request r) { _queue.push_back(std::move(r)); }

	// Called after retirements, at most once per `seconds` (jobs finished, jobs
	// submitted, seconds since run() started).
	void on_progress(std::function<void(size_t, size_t, double)> f, double seconds = 10) {
		_progress       = std::move(f);
		_progress_every = seconds;
	}

	// Runs until every submitted job has finished; returns one record per job in
	// submission order.
	std::vector<record> run() {
		std::vector<record> out(_queue.size());
		std::vector<size_t> order(_queue.size());  // jobs by arrival
		for (size_t i = 0; i < order.size(); i++) order[i] = i;
		std::stable_sort(order.begin(), order.end(),
		                 [&](size_t a, size_t b) { return _queue[a].arrival < _queue[b].arrival; });
		auto const t0 = std::chrono::steady_clock::now();
		auto now      = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
		size_t next   = 0;
		int64_t used  = 0;
		struct active {
			size_t job;
			snn* net;
			batch* b;
			Int left;
		};
		std::vector<active> act;
		size_t done        = 0;
		double last_report = 0;
		while (next < _queue.size() || !act.empty()) {
			// admit every job that has arrived and fits
			while (next < _queue.size() && _queue[order[next]].arrival <= now() &&
			       (act.empty() || used + _queue[order[next]].footprint <= _budget)) {
				request& r    = _queue[order[next]];
				batch* b      = r.single_step ? _single.get() : _multi.get();
				double const tb = now();
				snn& net      = b->add(r.dt, r.max_delay, r.seed);
				r.build(net);
				_t_build += now() - tb;
				used += r.footprint;
				out[order[next]] = {r.arrival, now(), 0.0, r.steps};
				act.push_back({order[next], &net, b, r.steps});
				next++;
			}
			if (act.empty()) {  // idle until the next arrival
				std::this_thread::sleep_for(std::chrono::duration<double>(
				    std::max(0.0, std::min(0.001, _queue[order[next]].arrival - now()))));
				continue;
			}
			// one round: every admitted job queues up to `window` steps; both
			// classes run their windows (asynchronously, on their own streams)
			double const tr = now();
			{
				struct held {  // every job's steps of the round go into one window per class
					batch* b[2];
					held(batch* m, batch* s) : b{m, s} {
						for (auto* x : b) x->hold(true);
					}
					~held() {
						for (auto* x : b) x->hold(false);
					}
				} h(_multi.get(), _single.get());
				for (auto& a : act) {
					Int const n = std::min(a.left, _window);
					for (Int k = 0; k < n; k++) a.net->step();
					a.left -= n;
				}
			}
			_multi->flush();
			_single->flush();
			_t_round += now() - tr;
			double const tt = now();
			// retire finished jobs
			for (size_t i = 0; i < act.size();) {
				if (act[i].left > 0) {
					i++;
					continue;
				}
				request& r = _queue[act[i].job];
				if (r.finish) r.finish(*act[i].net);
				out[act[i].job].finished = now();
				act[i].b->remove(*act[i].net);
				used -= r.footprint;
				act.erase(act.begin() + long(i));
				done++;
			}
			_t_retire += now() - tt;
			if (_progress && now() - last_report >= _progress_every) {
				last_report = now();
				_progress(done, _queue.size(), last_report);
			}
		}
		_queue.clear();
		return out;
	}

	// Synaptic events delivered by all jobs so far.
	int64_t events() { return events(*_multi) + events(*_single); }

	// Host seconds spent building admitted networks, launching rounds (including
	// waits on the device), and reading and freeing finished jobs.
	double build_seconds() const { return _t_build; }
	double round_seconds() const { return _t_round; }
	double retire_seconds() const { return _t_retire; }
	// Host seconds spent generating admitted jobs' connectivity (inside rounds).
	double generation_seconds() { return (csr_ms(*_multi) + csr_ms(*_single)) / 1e3; }

private:
	int64_t _budget;
	Int _window;
	double _t_build = 0, _t_round = 0, _t_retire = 0;
	std::unique_ptr<batch> _multi, _single;
	std::vector<request> _queue;
	std::function<void(size_t, size_t, double)> _progress;
	double _progress_every = 10;

	static double csr_ms(batch& b) {
		spice_stats s{};
		detail::check_spice(spice_ctx_stats(b.context(), &s));
		return s.csr_ms;
	}
	static int64_t events(batch& b) {
		spice_stats s{};
		detail::check_spice(spice_ctx_stats(b.context(), &s));
		return int64_t(s.events);
	}
};
}  // namespace spice
