// Agent container of a microenvironment (reference: reactions-diffusion/biofvm/include/biofvm/agent_container.h:8-14,
// common/include/common/generic_agent_container.h:48-103): owns the base and biofvm data, hands out
// agent views, removes by moving the last agent into the hole.
#pragma once

#include <memory>
#include <vector>

#include <common/agents.h>

#include "agent.h"

namespace physicore::biofvm {

using agent_container_interface = physicore::generic_agent_interface_container<agent_interface>;
using container_ptr = std::shared_ptr<agent_container_interface>;

class agent_container : public physicore::generic_agent_impl_container<agent>
{
	std::unique_ptr<physicore::base_agent_data> base;
	std::unique_ptr<agent_data> data;
	std::vector<std::unique_ptr<agent>> agents;

public:
	agent_container(std::unique_ptr<physicore::base_agent_data> base, std::unique_ptr<agent_data> data)
		: base(std::move(base)), data(std::move(data))
	{}

	agent* create() override
	{
		base->add();
		data->add();
		agents.emplace_back(std::make_unique<agent>(agents.size(), *data));
		return agents.back().get();
	}

	agent* get_agent_at(index_t position) override
	{
		return position < agents.size() ? agents[position].get() : nullptr;
	}

	void remove_agent(base_agent_interface* a) override { remove_at(a->get_index()); }

	void remove_at(index_t position) override
	{
		if (position >= agents.size())
			return;
		base->remove_at(position);
		data->remove_at(position);
		agents.back()->set_index(position);
		std::swap(agents[position], agents.back());
		agents.pop_back();
	}

	std::size_t size() const override { return agents.size(); }

	agent_data& agent_data_ref() override { return *data; }
	physicore::base_agent_data& base_data_ref() { return *base; }
};

} // namespace physicore::biofvm
