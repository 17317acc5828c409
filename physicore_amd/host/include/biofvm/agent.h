// A view of one agent over agent_data (reference: reactions-diffusion/biofvm/include/biofvm/agent.h,
// agent_generic_storage.h:15-76).
#pragma once

#include "agent_data.h"
#include "agent_interface.h"

namespace physicore::biofvm {

class agent : public agent_interface
{
	agent_data& data;

	std::span<real_t> row(std::vector<real_t>& v)
	{
		return std::span<real_t>(&v[index * data.substrate_count], data.substrate_count);
	}

public:
	using DataType = agent_data;
	using InterfaceType = agent_interface;

	agent(index_t index, agent_data& data) : base_agent_interface(index), data(data) {}

	std::span<real_t> position() override
	{
		return std::span<real_t>(&data.base_data.positions[index * data.base_data.dims], data.base_data.dims);
	}
	std::span<real_t> secretion_rates() override { return row(data.secretion_rates); }
	std::span<real_t> saturation_densities() override { return row(data.saturation_densities); }
	std::span<real_t> uptake_rates() override { return row(data.uptake_rates); }
	std::span<real_t> net_export_rates() override { return row(data.net_export_rates); }
	std::span<real_t> internalized_substrates() override { return row(data.internalized_substrates); }
	std::span<real_t> fraction_released_at_death() override { return row(data.fraction_released_at_death); }
	std::span<real_t> fraction_transferred_when_ingested() override
	{
		return row(data.fraction_transferred_when_ingested);
	}
	real_t& volume() override { return data.volumes[index]; }
	uint8_t& is_active() override { return data.is_active[index]; }
};

} // namespace physicore::biofvm
