// Agent container of a microenvironment (reference: reactions-diffusion/biofvm/include/biofvm/agent_container.h:8-14,
// common/include/common/generic_agent_container.h:48-103): owns the base and biofvm data, hands out
// agent views, removes by moving the last agent into the hole.  One concrete
// container for the one agent type of this module, where the reference
// composes one from a list of agent types.
#pragma once

#include <memory>
#include <vector>

#include <common/base_agent_data.h>
#include <common/generic_agent_container.h>

#include "agent.h"

namespace physicore::biofvm {

using agent_container_interface = physicore::generic_agent_interface_container<agent_interface>;
using container_ptr = std::shared_ptr<agent_container_interface>;

namespace detail {
// the arrays, constructed before the container base that refers to them
struct agent_arrays
{
	std::unique_ptr<physicore::base_agent_data> base;
	std::unique_ptr<agent_data> data;
};
} // namespace detail

class agent_container : private detail::agent_arrays, public physicore::generic_agent_impl_container<agent>
{
	std::vector<std::unique_ptr<agent>> agents;

public:
	agent_container(std::unique_ptr<physicore::base_agent_data> base_, std::unique_ptr<agent_data> data_)
		: detail::agent_arrays { std::move(base_), std::move(data_) },
		  physicore::generic_agent_impl_container<agent>(*detail::agent_arrays::data)
	{}

	agent* create() override
	{
		base->add();
		detail::agent_arrays::data->add();
		agents.emplace_back(std::make_unique<agent>(agents.size(), *detail::agent_arrays::data));
		return agents.back().get();
	}

	agent* get_agent_at(index_t position) override
	{
		return position < agents.size() ? agents[position].get() : nullptr;
	}

	void remove_agent(base_agent_interface* a) override { remove_at(get_agent_index(a)); }

	void remove_at(index_t position) override
	{
		if (position >= agents.size())
			return;
		base->remove_at(position);
		detail::agent_arrays::data->remove_at(position);
		get_agent_index(agents.back().get()) = position;
		std::swap(agents[position], agents.back());
		agents.pop_back();
	}

	std::size_t size() const override { return agents.size(); }
};

} // namespace physicore::biofvm
