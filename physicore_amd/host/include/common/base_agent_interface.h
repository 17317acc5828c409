// Per-agent view base of the agent containers (reference:
// common/include/common/base_agent_interface.h:13-25): an agent is an index
// into the structure-of-arrays data its container owns.
#pragma once

#include <span>

#include "types.h"

namespace physicore {

template <typename InterfaceType>
class generic_agent_interface_container;

class base_agent_interface
{
protected:
	// containers renumber their agents (removal moves the last agent into the hole)
	template <typename>
	friend class generic_agent_interface_container;

	index_t index;

public:
	explicit base_agent_interface(index_t index) : index(index) {}

	virtual std::span<real_t> position() = 0;
	virtual ~base_agent_interface() = default;
};

} // namespace physicore
