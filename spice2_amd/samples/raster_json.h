// Spike raster as JSON on stdout, byte-compatible with the printer the
// reference samples use without matplot (samples/matplot.cpp:111-157):
//   {\n\t"name": "<model>",\n\t"spikes": [\n\t\t[ids of step 0],\n\t\t[...]\n\t]\n}\n
// Ids of the populations streamed within one step are offset by the sizes of
// the populations streamed before them.
#pragma once

#include <iostream>
#include <string>

#include "spice/detail/neuron_population.h"

class raster_json {
public:
	explicit raster_json(std::string const& model, bool skip_silent_steps = false) :
	_skip(skip_silent_steps) {
		std::cout << "{\n\t\"name\": \"" << model << "\",\n\t\"spikes\": [\n";
	}
	~raster_json() { std::cout << "\n\t]\n}\n"; }

	raster_json& operator<<(spice::detail::NeuronPopulation const* pop) {
		auto const s = pop->spikes(0);
		if (!_skip || !s.empty()) {
			for (auto id : s) {
				std::cout << _sep << id + _offset;
				_sep = ",";
			}
			_offset += pop->size();
		}
		return *this;
	}
	raster_json& operator<<(char c) {
		if (c == '\n') {
			if (_offset > 0) {
				std::cout << "]";
				_sep = ",\n\t\t[";
			}
			_offset = 0;
		}
		return *this;
	}

private:
	bool _skip;
	Int _offset      = 0;
	char const* _sep = "\t\t[";
};
