#include <biofvm/vtk_serializer.h>

#include <algorithm>
#include <array>
#include <charconv>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <bfvm_hip.h>
#include <hip_solver/hip_solver.h>

using namespace physicore;
using namespace physicore::biofvm;

namespace {

std::string num(double v)
{
	char buf[32];
	auto r = std::to_chars(buf, buf + sizeof(buf), v);
	return std::string(buf, r.ptr);
}

// One ImageData file: `value(v, s)` is substrate s at linear voxel v.
template <typename Value>
void write_vti(const std::filesystem::path& path, const std::array<long long, 6>& extent,
			   const std::array<double, 3>& spacing, const std::vector<std::string>& names, std::size_t V,
			   Value value)
{
	const std::size_t S = names.size();
	const std::uint64_t block = (std::uint64_t)V * sizeof(double);
	std::ostringstream h;
	std::string ext;
	for (int k = 0; k < 6; k++)
		ext += (k ? " " : "") + std::to_string(extent[k]);
	h << "<?xml version=\"1.0\"?>\n"
	  << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
	  << "  <ImageData WholeExtent=\"" << ext << "\" Origin=\"0 0 0\" Spacing=\"" << num(spacing[0]) << " "
	  << num(spacing[1]) << " " << num(spacing[2]) << "\" Direction=\"1 0 0 0 1 0 0 0 1\">\n"
	  << "  <Piece Extent=\"" << ext << "\">\n"
	  << "    <PointData>\n"
	  << "    </PointData>\n"
	  << "    <CellData>\n";
	for (std::size_t s = 0; s < S; s++)
	{
		double lo = V ? value(0, s) : 0.0, hi = lo;
		for (std::size_t v = 1; v < V; v++)
		{
			const double x = value(v, s);
			lo = std::min(lo, x);
			hi = std::max(hi, x);
		}
		h << "      <DataArray type=\"Float64\" Name=\"" << names[s] << "\" format=\"appended\" RangeMin=\""
		  << num(lo) << "\" RangeMax=\"" << num(hi) << "\" offset=\"" << s * (sizeof(std::uint64_t) + block)
		  << "\"/>\n";
	}
	h << "    </CellData>\n"
	  << "  </Piece>\n"
	  << "  </ImageData>\n"
	  << "  <AppendedData encoding=\"raw\">\n   _";
	std::ofstream out(path, std::ios::binary);
	if (!out)
		throw std::runtime_error("vtk_serializer: cannot open " + path.string());
	const std::string head = h.str();
	out.write(head.data(), (std::streamsize)head.size());
	std::vector<double> chunk(std::min<std::size_t>(V, 1 << 16));
	for (std::size_t s = 0; s < S; s++)
	{
		out.write(reinterpret_cast<const char*>(&block), sizeof(block));
		for (std::size_t v0 = 0; v0 < V; v0 += chunk.size())
		{
			const std::size_t n = std::min(chunk.size(), V - v0);
			for (std::size_t k = 0; k < n; k++)
				chunk[k] = value(v0 + k, s);
			out.write(reinterpret_cast<const char*>(chunk.data()), (std::streamsize)(n * sizeof(double)));
		}
	}
	const char tail[] = "\n  </AppendedData>\n</VTKFile>\n";
	out.write(tail, sizeof(tail) - 1);
	out.close();
	if (!out)
		throw std::runtime_error("vtk_serializer: write failed for " + path.string());
}

} // namespace

struct vtk_serializer::impl
{
	std::filesystem::path output_dir, vtks_dir;
	std::string pvd_contents;
	std::size_t iteration = 0;

	std::array<long long, 6> extent {};
	std::array<double, 3> spacing {};
	std::vector<std::string> names;
	std::size_t V = 0, S = 0;

	// asynchronous path (hip_solver): two pinned host buffers, a writer thread
	struct job
	{
		int buf;
		bfvm_ctx* ctx;
		int64_t ticket;
		std::filesystem::path path;
	};
	std::array<float*, 2> bufs { nullptr, nullptr };
	std::vector<int> free_bufs;
	std::deque<job> queue;
	int writing = 0;
	bool stop = false;
	std::exception_ptr error;
	std::mutex mu;
	std::condition_variable cv;
	std::thread writer;

	~impl()
	{
		{
			std::unique_lock<std::mutex> lk(mu);
			stop = true;
		}
		cv.notify_all();
		if (writer.joinable())
			writer.join();
		for (float* b : bufs)
			(void)bfvm_host_free(b);
	}

	void rethrow()
	{
		if (error)
		{
			auto e = error;
			error = nullptr;
			std::rethrow_exception(e);
		}
	}

	// vtk_serializer_base::append_to_pvd (vtk_serializer_base.cpp:22-37): the collection is
	// rewritten whole after each file
	void append_to_pvd(const std::filesystem::path& file, real_t current_time)
	{
		pvd_contents += "    <DataSet timestep=\"" + std::to_string(current_time) + "\" group=\"\" part=\"0\" file=\"" +
						std::filesystem::relative(file, output_dir).string() + "\" />\n";
		std::ofstream pvd(output_dir / "microenvironment.pvd");
		pvd << pvd_contents << "  </Collection>\n</VTKFile>";
	}

	void writer_loop()
	{
		for (;;)
		{
			job j;
			{
				std::unique_lock<std::mutex> lk(mu);
				cv.wait(lk, [&] { return stop || !queue.empty(); });
				if (queue.empty())
					return;
				j = std::move(queue.front());
				queue.pop_front();
				writing++;
			}
			try
			{
				if (bfvm_snapshot_wait(j.ctx, j.ticket) != BFVM_OK)
					throw std::runtime_error(std::string("vtk_serializer: snapshot: ") + bfvm_last_error(nullptr));
				const float* f = bufs[j.buf];
				const std::size_t SS = S;
				write_vti(j.path, extent, spacing, names, V,
						  [&](std::size_t v, std::size_t s) { return (double)f[v * SS + s]; });
			}
			catch (...)
			{
				std::unique_lock<std::mutex> lk(mu);
				if (!error)
					error = std::current_exception();
			}
			{
				std::unique_lock<std::mutex> lk(mu);
				writing--;
				free_bufs.push_back(j.buf);
			}
			cv.notify_all();
		}
	}
};

// vtk_serializer::vtk_serializer (vtk_serializer.cpp:14-52) and vtk_serializer_base (vtk_serializer_base.cpp:8-20)
vtk_serializer::vtk_serializer(std::string_view output_dir, microenvironment& m) : p(std::make_unique<impl>())
{
	p->output_dir = std::filesystem::path(output_dir);
	p->vtks_dir = p->output_dir / "vtk_microenvironment";
	p->pvd_contents = "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" "
					  "byte_order=\"LittleEndian\">\n  <Collection>\n";
	std::filesystem::create_directories(p->vtks_dir);

	const auto& mesh = m.mesh;
	for (int a = 0; a < 3; a++)
	{
		const long long start = mesh.bounding_box_mins[a] / (sindex_t)mesh.voxel_shape[a];
		p->extent[2 * a] = start;
		p->extent[2 * a + 1] = start + (long long)mesh.grid_shape[a];
		p->spacing[a] = (double)mesh.voxel_shape[a];
	}
	if (mesh.dims != 3)
	{
		p->extent[4] = p->extent[5] = 0;
		p->spacing[2] = 0;
	}
	p->names = m.substrates_names;
	p->names.resize(m.substrates_count);
	p->S = m.substrates_count;
	p->V = mesh.voxel_count();
}

vtk_serializer::~vtk_serializer()
{
	try
	{
		flush();
	}
	catch (...)
	{
	}
}

void vtk_serializer::flush()
{
	std::unique_lock<std::mutex> lk(p->mu);
	p->cv.wait(lk, [&] { return p->queue.empty() && p->writing == 0; });
	p->rethrow();
}

void vtk_serializer::serialize(const microenvironment& m, real_t current_time)
{
	std::ostringstream ss;
	ss << "microenvironment_" << std::setw(6) << std::setfill('0') << p->iteration << ".vti";
	const std::filesystem::path path = p->vtks_dir / ss.str();

	auto* hs = dynamic_cast<kernels::hip_solver::hip_solver*>(m.solver.get());
	bfvm_ctx* ctx = hs ? hs->snapshot_context() : nullptr;
	if (ctx)
	{
		std::unique_lock<std::mutex> lk(p->mu);
		p->rethrow();
		if (!p->bufs[0])
		{
			for (float*& b : p->bufs)
			{
				void* q = nullptr;
				if (bfvm_host_alloc(p->V * p->S * sizeof(float), &q) != BFVM_OK)
					throw std::runtime_error(std::string("vtk_serializer: ") + bfvm_last_error(nullptr));
				b = static_cast<float*>(q);
			}
			p->free_bufs = { 0, 1 };
			p->writer = std::thread([this] { p->writer_loop(); });
		}
		p->cv.wait(lk, [&] { return !p->free_bufs.empty(); });
		const int b = p->free_bufs.back();
		p->free_bufs.pop_back();
		int64_t ticket = 0;
		if (bfvm_snapshot_begin(ctx, p->bufs[b], &ticket) != BFVM_OK)
		{
			p->free_bufs.push_back(b);
			throw std::runtime_error(std::string("vtk_serializer: snapshot: ") + bfvm_last_error(ctx));
		}
		p->queue.push_back({ b, ctx, ticket, path });
		lk.unlock();
		p->cv.notify_all();
	}
	else
	{
		// any other solver: the reference's per-voxel read (vtk_serializer.cpp:56-64)
		const auto& mesh = m.mesh;
		const std::size_t S = p->S;
		std::vector<double> f(p->V * S);
		for (index_t z = 0; z < mesh.grid_shape[2]; ++z)
			for (index_t y = 0; y < mesh.grid_shape[1]; ++y)
				for (index_t x = 0; x < mesh.grid_shape[0]; ++x)
					for (index_t s = 0; s < S; ++s)
						f[mesh.linearize(x, y, z) * S + s] = m.get_substrate_density(s, x, y, z);
		write_vti(path, p->extent, p->spacing, p->names, p->V,
				  [&](std::size_t v, std::size_t s) { return f[v * S + s]; });
	}
	p->append_to_pvd(path, current_time);
	p->iteration++;
}
