// Agent storage behind the biofvm interface: structure-of-arrays data, a
// container that hands out per-agent views, and the accessor a solver uses to
// reach the arrays.  Same names and members as the reference
// (common/include/common/base_agent_data_generic_storage.h:33-78,
// generic_agent_container.h:13-103, generic_agent_solver.h:13-23), with a
// single concrete container instead of the reference's variadic template.
#pragma once

#include <cassert>
#include <cstring>
#include <memory>
#include <span>
#include <vector>

#include "types.h"

namespace physicore {

struct base_agent_data
{
	index_t agents_count = 0;
	index_t dims;
	std::vector<real_t> positions; // [agents_count * dims]

	explicit base_agent_data(index_t dims = 3) : dims(dims) {}

	void add()
	{
		++agents_count;
		positions.resize(agents_count * dims);
	}

	// removal moves the last agent into the hole (generic_agent_container.h:78-91)
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

class base_agent_interface
{
protected:
	index_t index;

public:
	explicit base_agent_interface(index_t index) : index(index) {}
	index_t get_index() const { return index; }
	void set_index(index_t i) { index = i; }
	virtual std::span<real_t> position() = 0;
	virtual ~base_agent_interface() = default;
};

template <typename InterfaceType>
class generic_agent_interface_container
{
public:
	virtual InterfaceType* create() = 0;
	virtual InterfaceType* get_agent_at(index_t position) = 0;
	virtual void remove_agent(base_agent_interface* agent) = 0;
	virtual void remove_at(index_t position) = 0;
	virtual std::size_t size() const = 0;
	virtual ~generic_agent_interface_container() = default;
};

// A container that owns the data of AgentType and exposes it to solvers.
template <typename AgentType>
class generic_agent_impl_container : public generic_agent_interface_container<typename AgentType::InterfaceType>
{
public:
	virtual typename AgentType::DataType& agent_data_ref() = 0;
};

template <typename AgentType>
class generic_agent_solver
{
public:
	typename AgentType::DataType& retrieve_agent_data(
		generic_agent_interface_container<typename AgentType::InterfaceType>& container)
	{
		return dynamic_cast<generic_agent_impl_container<AgentType>&>(container).agent_data_ref();
	}
};

} // namespace physicore
