#include <biofvm/mesh.h>

#include <cassert>

using namespace physicore;
using namespace physicore::biofvm;

cartesian_mesh::cartesian_mesh(index_t dims, std::array<sindex_t, 3> mins, std::array<sindex_t, 3> maxs,
							   std::array<index_t, 3> shape)
	: dims(dims), bounding_box_mins(mins), bounding_box_maxs(maxs), voxel_shape(shape), grid_shape({ 1, 1, 1 })
{
	// ceil((max - min) / d) on the axes below dims (reference mesh.cpp:8-28)
	for (index_t a = 0; a < dims && a < 3; a++)
		grid_shape[a] = (index_t)((maxs[a] - mins[a] + (sindex_t)shape[a] - 1) / (sindex_t)shape[a]);
}

std::size_t cartesian_mesh::voxel_count() const
{
	return (std::size_t)grid_shape[0] * (std::size_t)grid_shape[1] * (std::size_t)grid_shape[2];
}

index_t cartesian_mesh::voxel_volume() const { return voxel_shape[0] * voxel_shape[1] * voxel_shape[2]; }

std::size_t cartesian_mesh::linearize(index_t x, index_t y, index_t z) const
{
	return x + y * grid_shape[0] + z * grid_shape[0] * grid_shape[1];
}

std::array<index_t, 3> cartesian_mesh::voxel_position(std::span<const real_t> position) const
{
	// truncation of (p - min) / d in FP64 (reference mesh.cpp:37-61)
	assert(position.size() == (std::size_t)dims);
	std::array<index_t, 3> v = { 0, 0, 0 };
	for (std::size_t a = 0; a < position.size() && a < 3; a++)
		v[a] = (index_t)((position[a] - (real_t)bounding_box_mins[a]) / (real_t)voxel_shape[a]);
	return v;
}

std::array<real_t, 3> cartesian_mesh::voxel_center(std::array<index_t, 3> p) const
{
	std::array<real_t, 3> c;
	for (int a = 0; a < 3; a++)
		c[a] = (real_t)(p[a] * voxel_shape[a] + bounding_box_mins[a]) + (real_t)voxel_shape[a] / (real_t)2.0;
	return c;
}
