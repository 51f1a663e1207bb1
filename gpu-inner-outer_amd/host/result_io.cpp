// result_io.cpp — HDF5 result files with the reference's layouts: a GPU run
// (src/result_io.cpp:364-517) and the --cpu / --compare runs
// (src/result_io.cpp:238-362). Contiguous, uncompressed, native-endian
// datasets; fixed-length NULLTERM strings; readers reshape with order="F"
// (x fastest).
#include "gsc/result_io.hpp"

#include <cctype>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

#if GSC_HAS_HDF5
#include <hdf5.h>
#endif

namespace gsc {

namespace {

std::string sanitize(const std::string& in) {
    std::string out;
    for (unsigned char ch : in) out.push_back((std::isalnum(ch) || ch == '_' || ch == '-') ? static_cast<char>(ch) : '_');
    return out.empty() ? std::string("unnamed") : out;
}

#if GSC_HAS_HDF5

// Owning hid_t with its close function.
class Id {
public:
    Id(hid_t id, herr_t (*close)(hid_t), const std::string& what) : id_(id), close_(close) {
        if (id_ < 0) throw std::runtime_error("HDF5: " + what);
    }
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    ~Id() {
        if (id_ >= 0) close_(id_);
    }
    hid_t get() const { return id_; }

private:
    hid_t id_;
    herr_t (*close_)(hid_t);
};

template <typename T>
hid_t native();
template <>
hid_t native<double>() { return H5T_NATIVE_DOUBLE; }
template <>
hid_t native<std::uint8_t>() { return H5T_NATIVE_UINT8; }
template <>
hid_t native<std::uint32_t>() { return H5T_NATIVE_UINT32; }
template <>
hid_t native<std::uint64_t>() { return H5T_NATIVE_UINT64; }
template <>
hid_t native<int>() { return H5T_NATIVE_INT; }

void check(herr_t e, const std::string& what) {
    if (e < 0) throw std::runtime_error("HDF5: " + what);
}

Id group(hid_t parent, const char* name) {
    return Id(H5Gcreate2(parent, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), &H5Gclose,
              std::string("cannot create group ") + name);
}

template <typename T>
void write_dataset(hid_t parent, const char* name, const T* data, int rank, const hsize_t* dims) {
    Id space(rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, dims, nullptr), &H5Sclose,
             std::string("dataspace ") + name);
    Id ds(H5Dcreate2(parent, name, native<T>(), space.get(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), &H5Dclose,
          std::string("dataset ") + name);
    bool empty = false;
    for (int k = 0; k < rank; ++k) empty = empty || dims[k] == 0;
    if (!empty) check(H5Dwrite(ds.get(), native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data), std::string("write ") + name);
}

template <typename T>
void scalar(hid_t parent, const char* name, T value) {
    write_dataset(parent, name, &value, 0, nullptr);
}

template <typename T>
void vec(hid_t parent, const char* name, const T* data, std::size_t n) {
    const hsize_t dims[1] = {static_cast<hsize_t>(n)};
    write_dataset(parent, name, data, 1, dims);
}

template <typename T>
void vec(hid_t parent, const char* name, const std::vector<T>& v) {
    vec(parent, name, v.data(), v.size());
}

template <typename T>
void matrix(hid_t parent, const char* name, const T* data, std::size_t rows, std::size_t cols) {
    const hsize_t dims[2] = {static_cast<hsize_t>(rows), static_cast<hsize_t>(cols)};
    write_dataset(parent, name, data, 2, dims);
}

void string(hid_t parent, const char* name, const std::string& value) {
    Id type(H5Tcopy(H5T_C_S1), &H5Tclose, std::string("string type ") + name);
    check(H5Tset_size(type.get(), std::max<std::size_t>(1, value.size() + 1)), "string size");
    check(H5Tset_strpad(type.get(), H5T_STR_NULLTERM), "string pad");
    Id space(H5Screate(H5S_SCALAR), &H5Sclose, "string dataspace");
    Id ds(H5Dcreate2(parent, name, type.get(), space.get(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), &H5Dclose,
          std::string("dataset ") + name);
    check(H5Dwrite(ds.get(), type.get(), H5S_ALL, H5S_ALL, H5P_DEFAULT, value.c_str()), std::string("write ") + name);
}

// /meta, /state, /input and /config (result_io.cpp:238-290, 380-440).
void write_header(hid_t f, const CaseConfig& cfg, const Grid4D<std::uint32_t>& sg, const InputGrid2D& ig,
                  std::uint32_t run_count) {
    {
        Id meta = group(f, "meta");
        string(meta.get(), "case_name", cfg.name);
        scalar<std::uint32_t>(meta.get(), "invalid_input", kInvalidInput);
        scalar<std::uint32_t>(meta.get(), "unreachable_step", kUnreachableStep);
        scalar<std::uint32_t>(meta.get(), "run_count", run_count);
    }
    {
        Id st = group(f, "state");
        vec(st.get(), "lb", sg.lb, kStateDim);
        vec(st.get(), "ub", sg.ub, kStateDim);
        vec(st.get(), "eta", sg.eta, kStateDim);
        vec(st.get(), "shape", sg.shape, kStateDim);
        vec(st.get(), "strides", sg.strides, kStateDim);
        scalar<std::uint64_t>(st.get(), "total_size", sg.total_size);
        scalar<int>(st.get(), "wrap_dim", sg.wrap_dim);
    }
    {
        Id in = group(f, "input");
        vec(in.get(), "lb", ig.lb, kInputDim);
        vec(in.get(), "ub", ig.ub, kInputDim);
        vec(in.get(), "eta", ig.eta, kInputDim);
        vec(in.get(), "shape", ig.shape, kInputDim);
        vec(in.get(), "strides", ig.strides, kInputDim);
        scalar<std::uint64_t>(in.get(), "total_size", ig.total_size);
        std::vector<double> centers(ig.total_size * kInputDim);
        for (std::uint64_t k = 0; k < ig.total_size; ++k)
            ig.center(static_cast<InputIndex>(k), &centers[k * kInputDim]);
        matrix(in.get(), "centers", centers.data(), ig.total_size, kInputDim);
    }
    {
        Id cf = group(f, "config");
        vec(cf.get(), "candidate_lb", cfg.candidate.lb, kStateDim);
        vec(cf.get(), "candidate_ub", cfg.candidate.ub, kStateDim);
        vec(cf.get(), "map_lb", cfg.map.lb, kStateDim);
        vec(cf.get(), "map_ub", cfg.map.ub, kStateDim);
        vec(cf.get(), "disturbance_half_width", cfg.disturbance_half_width, kStateDim);
        scalar<double>(cf.get(), "sample_time", cfg.sample_time);
        scalar<std::uint32_t>(cf.get(), "integration_substeps", cfg.integration_substeps);
        scalar<std::uint32_t>(cf.get(), "max_iterations", cfg.max_iterations);
        scalar<std::uint8_t>(cf.get(), "verbose", cfg.verbose ? 1 : 0);
        std::vector<double> obst;
        for (const auto& o : cfg.obstacles) {
            obst.insert(obst.end(), o.lb, o.lb + kStateDim);
            obst.insert(obst.end(), o.ub, o.ub + kStateDim);
        }
        matrix(cf.get(), "obstacles", obst.data(), cfg.obstacles.size(), 2 * kStateDim);
    }
}

// /summary/<label> and /result/<label> of one run (result_io.cpp:292-315).
void write_run(hid_t summary, hid_t result, const std::string& label, const SolveResult<std::uint32_t>& r,
               double abstraction_ms, double solve_ms, const std::uint64_t* pair_count) {
    const std::string name = sanitize(label);
    {
        Id g = group(summary, name.c_str());
        scalar<std::uint32_t>(g.get(), "iterations", r.iterations);
        scalar<std::uint8_t>(g.get(), "converged", r.converged ? 1 : 0);
        string(g.get(), "message", r.message);
        scalar<std::uint64_t>(g.get(), "reachable_states", r.reachable_states);
        scalar<std::uint64_t>(g.get(), "certified_candidate_states", r.certified_candidate_states);
        scalar<double>(g.get(), "abstraction_ms", abstraction_ms);
        scalar<double>(g.get(), "solve_ms", solve_ms);
        if (pair_count) scalar<std::uint64_t>(g.get(), "pair_count", *pair_count);
    }
    {
        Id g = group(result, name.c_str());
        vec(g.get(), "valid_mask", r.valid_mask);
        vec(g.get(), "candidate_mask", r.candidate_mask);
        vec(g.get(), "reachable_mask", r.reachable_mask);
        vec(g.get(), "controller", r.controller);
        vec(g.get(), "reach_step", r.reach_step);
    }
}

Id create_file(const std::string& output_path) {
    const std::filesystem::path out(output_path);
    if (out.has_parent_path()) std::filesystem::create_directories(out.parent_path());
    return Id(H5Fcreate(out.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), &H5Fclose,
              "cannot create file " + out.string());
}

#endif

}  // namespace

bool hdf5_export_compiled() {
#if GSC_HAS_HDF5
    return true;
#else
    return false;
#endif
}

std::string default_result_path(const CaseConfig& cfg, const std::string& mode) {
    return (std::filesystem::path("results") / (sanitize(cfg.name) + "_" + sanitize(mode) + ".h5")).string();
}

void write_results_hdf5(const std::string& output_path, const CaseConfig& cfg, const Grid4D<std::uint32_t>& state_grid,
                        const InputGrid2D& input_grid, double abstraction_ms,
                        const std::vector<ResultExportView>& runs) {
#if GSC_HAS_HDF5
    if (runs.empty()) throw std::runtime_error("no runs were provided for HDF5 export");
    Id file = create_file(output_path);
    write_header(file.get(), cfg, state_grid, input_grid, static_cast<std::uint32_t>(runs.size()));
    Id summary = group(file.get(), "summary");
    Id result = group(file.get(), "result");
    for (const auto& run : runs) {
        if (!run.result) throw std::runtime_error("result export view has null result pointer");
        write_run(summary.get(), result.get(), run.label, *run.result, abstraction_ms, run.result->solve_ms, nullptr);
    }
#else
    (void)output_path;
    (void)cfg;
    (void)state_grid;
    (void)input_grid;
    (void)abstraction_ms;
    (void)runs;
    throw std::runtime_error("HDF5 export support was not compiled");
#endif
}

void write_gpu_results_hdf5(const std::string& output_path, const CaseConfig& cfg, const GpuRunReport& r) {
#if GSC_HAS_HDF5
    if (!r.executed) throw std::runtime_error("GPU report has not been executed");
    Id file = create_file(output_path);
    const hid_t f = file.get();
    write_header(f, cfg, r.state_grid, r.input_grid, 1);
    {
        Id summary = group(f, "summary");
        Id result = group(f, "result");
        write_run(summary.get(), result.get(), "gpu", r.result, r.abstraction_ms, r.solve_ms, &r.pair_count);
    }
    {
        Id a = group(f, "abstraction");
        scalar<std::uint64_t>(a.get(), "pair_count", r.pair_count);
        vec(a.get(), "min_flat", r.abstraction_min_flat);
        vec(a.get(), "max_flat", r.abstraction_max_flat);
        vec(a.get(), "valid", r.abstraction_valid);
    }
    {
        Id it = group(f, "iterations");
        if (!r.iteration_stats.empty()) {
            const std::size_t n = r.iteration_stats.size();
            std::vector<std::uint32_t> num(n);
            std::vector<std::uint64_t> nr(n), nc(n), tr(n), tc(n);
            std::vector<double> ims(n), pms(n), sms(n), rms(n);
            for (std::size_t k = 0; k < n; ++k) {
                const auto& s = r.iteration_stats[k];
                num[k] = s.iteration;
                nr[k] = s.newly_reachable;
                nc[k] = s.newly_certified;
                tr[k] = s.total_reachable;
                tc[k] = s.total_certified;
                ims[k] = s.iteration_ms;
                pms[k] = s.prefix_build_ms;
                sms[k] = s.satisfaction_check_ms;
                rms[k] = s.reduction_ms;
            }
            vec(it.get(), "iteration", num);
            vec(it.get(), "newly_reachable", nr);
            vec(it.get(), "newly_certified", nc);
            vec(it.get(), "total_reachable", tr);
            vec(it.get(), "total_certified", tc);
            vec(it.get(), "iteration_ms", ims);
            vec(it.get(), "prefix_build_ms", pms);
            vec(it.get(), "satisfaction_check_ms", sms);
            vec(it.get(), "reduction_ms", rms);
        }
    }
    {
        Id t = group(f, "timings");
        const KernelTimings& k = r.kernel_timings;
        scalar<double>(t.get(), "abstraction_kernel_ms", k.abstraction_kernel_ms);
        scalar<double>(t.get(), "abstraction_memcpy_h2d_ms", k.abstraction_memcpy_h2d_ms);
        scalar<double>(t.get(), "scatter_mask_ms", k.scatter_mask_ms);
        scalar<double>(t.get(), "scan_axis0_ms", k.scan_axis0_ms);
        scalar<double>(t.get(), "scan_axis1_ms", k.scan_axis1_ms);
        scalar<double>(t.get(), "scan_axis2_ms", k.scan_axis2_ms);
        scalar<double>(t.get(), "scan_axis3_ms", k.scan_axis3_ms);
        scalar<double>(t.get(), "prefix_build_total_ms", k.prefix_build_total_ms);
        scalar<double>(t.get(), "pair_satisfaction_ms", k.pair_satisfaction_ms);
        scalar<double>(t.get(), "reduce_inputs_ms", k.reduce_inputs_ms);
        scalar<double>(t.get(), "iteration_memcpy_ms", k.iteration_memcpy_ms);
        scalar<double>(t.get(), "result_memcpy_d2h_ms", k.result_memcpy_d2h_ms);
        scalar<double>(t.get(), "total_kernel_ms", k.total_kernel_ms);
        scalar<double>(t.get(), "total_memcpy_ms", k.total_memcpy_ms);
        scalar<double>(t.get(), "total_compute_ms", k.total_compute_ms);
    }
#else
    (void)output_path;
    (void)cfg;
    (void)r;
    throw std::runtime_error("HDF5 export support was not compiled");
#endif
}

}  // namespace gsc
