#include <biofvm/agent_data.h>

#include <cassert>
#include <cstring>

using namespace physicore;
using namespace physicore::biofvm;

void agent_data::add()
{
	++agents_count;
	const std::size_t n = agents_count * substrate_count;
	secretion_rates.resize(n);
	saturation_densities.resize(n);
	uptake_rates.resize(n);
	net_export_rates.resize(n);
	internalized_substrates.resize(n);
	fraction_released_at_death.resize(n);
	fraction_transferred_when_ingested.resize(n, 1);
	volumes.resize(agents_count);
	is_active.resize(agents_count, 1);
}

void agent_data::remove_at(index_t position)
{
	assert(position < agents_count);
	if (position >= agents_count)
		return;
	--agents_count;
	if (position < agents_count)
	{
		auto move_row = [&](std::vector<real_t>& v) {
			std::memcpy(&v[position * substrate_count], &v[agents_count * substrate_count],
						substrate_count * sizeof(real_t));
		};
		move_row(secretion_rates);
		move_row(saturation_densities);
		move_row(uptake_rates);
		move_row(net_export_rates);
		move_row(internalized_substrates);
		move_row(fraction_released_at_death);
		move_row(fraction_transferred_when_ingested);
		volumes[position] = volumes[agents_count];
		is_active[position] = is_active[agents_count];
	}
	const std::size_t n = agents_count * substrate_count;
	secretion_rates.resize(n);
	saturation_densities.resize(n);
	uptake_rates.resize(n);
	net_export_rates.resize(n);
	internalized_substrates.resize(n);
	fraction_released_at_death.resize(n);
	fraction_transferred_when_ingested.resize(n);
	volumes.resize(agents_count);
	is_active.resize(agents_count);
}
