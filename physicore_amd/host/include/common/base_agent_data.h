// Positions of every agent, [agents_count * dims] (reference:
// common/include/common/base_agent_data.h, base_agent_data_generic_storage.h:21-67).
#pragma once

#include <cassert>
#include <cstring>
#include <vector>

#include "types.h"

namespace physicore {

struct base_agent_data
{
	index_t agents_count = 0;
	index_t dims;
	std::vector<real_t> positions;

	explicit base_agent_data(index_t dims = 3) : dims(dims) {}

	void add()
	{
		++agents_count;
		positions.resize(agents_count * dims);
	}

	// the last agent moves into the hole (generic_agent_container.h:78-91)
	void remove_at(index_t position)
	{
		assert(position < agents_count);
		if (position >= agents_count)
			return;
		--agents_count;
		if (position != agents_count)
			std::memcpy(&positions[position * dims], &positions[agents_count * dims], dims * sizeof(real_t));
		positions.resize(agents_count * dims);
	}
};

} // namespace physicore
