// How a solver reaches the agent arrays of a microenvironment's container
// (reference: common/include/common/generic_agent_solver.h:13-23).
#pragma once

#include "generic_agent_container.h"

namespace physicore {

template <typename AgentType>
class generic_agent_solver
{
public:
	typename AgentType::DataType& retrieve_agent_data(
		generic_agent_interface_container<typename AgentType::InterfaceType>& container)
	{
		return dynamic_cast<generic_agent_impl_container<AgentType>&>(container).data;
	}
};

} // namespace physicore
