// Per-agent arrays of the biofvm module (reference:
// reactions-diffusion/biofvm/include/biofvm/agent_data_generic_storage.h:10-112): FP64, n*S row-major.
#pragma once

#include <vector>

#include <common/base_agent_data.h>
#include <common/types.h>

namespace physicore::biofvm {

struct agent_data
{
	physicore::base_agent_data& base_data;

	// n * substrate_count
	std::vector<real_t> secretion_rates;
	std::vector<real_t> saturation_densities;
	std::vector<real_t> uptake_rates;
	std::vector<real_t> net_export_rates;

	std::vector<real_t> internalized_substrates;
	std::vector<real_t> fraction_released_at_death;
	std::vector<real_t> fraction_transferred_when_ingested;

	// n
	std::vector<real_t> volumes;
	std::vector<uint8_t> is_active; // 1 active, 0 inactive

	index_t agents_count = 0;
	index_t substrate_count;

	explicit agent_data(physicore::base_agent_data& base_data, index_t substrate_count = 1)
		: base_data(base_data), substrate_count(substrate_count)
	{}

	void add();
	void remove_at(index_t position);
};

} // namespace physicore::biofvm
