// Per-agent view (reference: reactions-diffusion/biofvm/include/biofvm/agent_interface.h:10-30).
#pragma once

#include <span>

#include <common/base_agent_interface.h>
#include <common/types.h>

namespace physicore::biofvm {

class agent_interface : public virtual base_agent_interface
{
public:
	agent_interface() : base_agent_interface(0) {}
	virtual std::span<real_t> secretion_rates() = 0;
	virtual std::span<real_t> saturation_densities() = 0;
	virtual std::span<real_t> uptake_rates() = 0;
	virtual std::span<real_t> net_export_rates() = 0;
	virtual std::span<real_t> internalized_substrates() = 0;
	virtual std::span<real_t> fraction_released_at_death() = 0;
	virtual std::span<real_t> fraction_transferred_when_ingested() = 0;
	virtual real_t& volume() = 0;
	virtual uint8_t& is_active() = 0;
};

} // namespace physicore::biofvm
