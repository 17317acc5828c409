// Cartesian mesh of a microenvironment (reference: reactions-diffusion/biofvm/include/biofvm/mesh.h:11-32,
// src/mesh.cpp:8-77): grid shape ceil((max-min)/d) on axes below dims, 1 elsewhere.
#pragma once

#include <array>
#include <span>

#include <common/types.h>

namespace physicore::biofvm {

struct cartesian_mesh
{
	index_t dims; // 1, 2 or 3

	std::array<sindex_t, 3> bounding_box_mins;
	std::array<sindex_t, 3> bounding_box_maxs;

	std::array<index_t, 3> voxel_shape;
	std::array<index_t, 3> grid_shape;

	cartesian_mesh(index_t dims, std::array<sindex_t, 3> bounding_box_mins, std::array<sindex_t, 3> bounding_box_maxs,
				   std::array<index_t, 3> voxel_shape);

	std::size_t voxel_count() const;
	index_t voxel_volume() const;

	std::size_t linearize(index_t x, index_t y, index_t z) const;

	std::array<index_t, 3> voxel_position(std::span<const real_t> position) const;

	std::array<real_t, 3> voxel_center(std::array<index_t, 3> position) const;
};

} // namespace physicore::biofvm
