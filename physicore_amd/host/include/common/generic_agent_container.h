// Agent containers (reference: common/include/common/generic_agent_container.h:13-46):
// the interface a microenvironment holds, and the container of one agent type,
// whose data a solver reaches through generic_agent_solver.
#pragma once

#include "base_agent_interface.h"
#include "types.h"

namespace physicore {

template <typename InterfaceType>
class generic_agent_interface_container
{
protected:
	index_t& get_agent_index(base_agent_interface* agent) { return agent->index; }

public:
	virtual InterfaceType* create() = 0;
	virtual InterfaceType* get_agent_at(index_t position) = 0;
	virtual void remove_agent(base_agent_interface* agent) = 0;
	virtual void remove_at(index_t position) = 0;
	virtual std::size_t size() const = 0;
	virtual ~generic_agent_interface_container() = default;
};

template <typename AgentType>
class generic_agent_solver;

template <typename AgentType>
class generic_agent_impl_container : public generic_agent_interface_container<typename AgentType::InterfaceType>
{
protected:
	friend generic_agent_solver<AgentType>;

	typename AgentType::DataType& data;

public:
	explicit generic_agent_impl_container(typename AgentType::DataType& data) : data(data) {}
};

} // namespace physicore
