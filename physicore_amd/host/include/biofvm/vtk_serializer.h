// VTK output of the substrate densities (reference: reactions-diffusion/biofvm/src/vtk_serializer.h:14-26,
// vtk_serializer.cpp:14-82, vtk_serializer_base.cpp:8-37).
//
// Each serialize() writes output_dir/vtk_microenvironment/microenvironment_NNNNNN.vti, an
// ImageData file with one Float64 cell array per substrate, named after it
// (extent bounding_box_mins / voxel_shape + grid_shape, spacing voxel_shape; a
// 2-D or 1-D mesh gets a flat z extent and zero z spacing), and rewrites
// output_dir/microenvironment.pvd, the collection of every file with its time.
// The files are VTK XML written directly (raw appended data, UInt64 block
// headers), without the VTK library.
//
// With hip_solver, serialize() does not read the field voxel by voxel: it
// queues a device snapshot of the field after the steps solved so far, whose
// copy to pinned host memory runs while later steps do, and a writer thread
// converts the snapshot and writes the file.  serialize() returns once the
// snapshot is queued (it waits only while both host buffers are in use);
// flush() and the destructor wait for the files.  Other solvers are read
// through get_substrate_density, synchronously, as the reference does.
#pragma once

#include <memory>
#include <string_view>

#include "microenvironment.h"
#include "serializer.h"

namespace physicore::biofvm {

class vtk_serializer : public serializer
{
public:
	vtk_serializer(std::string_view output_dir, microenvironment& m);
	~vtk_serializer() override;

	vtk_serializer(const vtk_serializer&) = delete;
	vtk_serializer& operator=(const vtk_serializer&) = delete;

	void serialize(const microenvironment& m, real_t current_time) override;

	// Waits until every file queued so far is written; rethrows a writer error.
	void flush();

private:
	struct impl;
	std::unique_ptr<impl> p;
};

} // namespace physicore::biofvm
