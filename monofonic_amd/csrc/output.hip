// Particle output: the Gadget-HDF5 writer of the gadget_hdf5 /
// gadget_hdf5_double output plugins (output_gadget_hdf5.cc:36-318) and the
// SWIFT writer (output_swift.cc:36-377), fed straight from the device
// particle arrays of a run.
//
// File layout and datatypes follow the plugins and HDF_IO.hh:
//   /ICs_parameters  scalar attributes describing the run
//                    (output_gadget_hdf5.cc:140-170, output_swift.cc:135-177,
//                    HDF_IO.hh:1141-1262)
//   Gadget: /PartType1/Coordinates, /Velocities  [N][3], stored as 32-bit
//                    floats whatever the plugin's precision (HDF_IO.hh:850-885
//                    creates them with H5T_NATIVE_FLOAT);
//                    /PartType1/ParticleIDs [N] uint32 or (UseLongids) uint64;
//                    /Header 1-element attribute arrays
//                    (output_gadget_hdf5.cc:175-210)
//   SWIFT:  /Units (output_swift.cc:124-130); /PartType1/Coordinates,
//                    /Velocities [N][3] 64-bit floats, /ParticleIDs, /Masses
//                    [N] 64-bit, every dataset chunked with a Fletcher-32
//                    checksum (HDF_IO.hh:888-963); /Header (output_swift.cc:
//                    183-200)
//   IDs: id = 2^lattice * cell + sub-lattice (particle_generator.hh:229-251)
// The device converts each chunk of FP64 [p][3] values to the file's type
// and generates the IDs, so PCIe carries only the bytes written; two pinned
// staging buffers let the copy of chunk i+1 run while HDF5 writes chunk i.
//
// The HDF5 C library (1.10 API) is loaded at run time with dlopen, so the
// library builds and runs without it until a file is written: the path in
// MFIC_HDF5_LIB, then libhdf5.so on the loader path, then the conda copy.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cosmology.hpp"
#include "mfic_internal.hpp"
#include "output.hpp"

namespace mfic {

// ------------------------------------------------------------ kernels ----
// positions / velocities of particles [p0, p0 + cnt) as 32-bit floats
__global__ void k_stage_f32(const double* __restrict__ src, float* __restrict__ dst, size_t n3) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) dst[i] = (float)src[i];
}

// ParticleIDs of particles [p0, p0 + cnt): particle p = iload * ncell + cell
// has id id_offset + overload * cell + iload (particle_generator.hh:229-251)
template <typename I>
__global__ void k_stage_ids(I* __restrict__ dst, size_t p0, size_t cnt, size_t ncell, unsigned overload,
                            uint64_t id_offset) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const size_t p = p0 + i;
    const size_t iload = p / ncell, cell = p - iload * ncell;
    dst[i] = (I)(id_offset + overload * cell + iload);
}

// ------------------------------------------------------ HDF5 binding ----
// The handful of HDF5 1.10 C API entry points and constants the writer uses
// (hid_t is 64-bit since 1.10; H5P_DEFAULT, H5S_ALL = 0; H5F_ACC_TRUNC =
// 0x2; H5S_SCALAR = 0; H5S_SELECT_SET = 0; H5T_STR_NULLPAD = 1).
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

struct Hdf5 {
    void* so = nullptr;
    herr_t (*open)(void) = nullptr;
    hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t) = nullptr;
    herr_t (*Fclose)(hid_t) = nullptr;
    hid_t (*Gcreate2)(hid_t, const char*, hid_t, hid_t, hid_t) = nullptr;
    herr_t (*Gclose)(hid_t) = nullptr;
    hid_t (*Screate)(int) = nullptr;
    hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*) = nullptr;
    herr_t (*Sselect_hyperslab)(hid_t, int, const hsize_t*, const hsize_t*, const hsize_t*, const hsize_t*) = nullptr;
    herr_t (*Sclose)(hid_t) = nullptr;
    hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*) = nullptr;
    herr_t (*Dclose)(hid_t) = nullptr;
    hid_t (*Acreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t) = nullptr;
    herr_t (*Awrite)(hid_t, hid_t, const void*) = nullptr;
    herr_t (*Aclose)(hid_t) = nullptr;
    hid_t (*Tcopy)(hid_t) = nullptr;
    herr_t (*Tset_size)(hid_t, size_t) = nullptr;
    herr_t (*Tset_strpad)(hid_t, int) = nullptr;
    herr_t (*Tclose)(hid_t) = nullptr;
    herr_t (*Eset_auto2)(hid_t, void*, void*) = nullptr;
    hid_t (*Pcreate)(hid_t) = nullptr;
    herr_t (*Pset_chunk)(hid_t, int, const hsize_t*) = nullptr;
    herr_t (*Pset_fletcher32)(hid_t) = nullptr;
    herr_t (*Pclose)(hid_t) = nullptr;
    hid_t FLOAT = -1, DOUBLE = -1, INT = -1, UINT = -1, ULLONG = -1, C_S1 = -1, DATASET_CREATE = -1;
    std::string error;
};

template <typename F> static bool sym(void* so, const char* name, F& f) {
    f = reinterpret_cast<F>(dlsym(so, name));
    return f != nullptr;
}

static Hdf5& hdf5() {
    static Hdf5 h;
    static std::once_flag once;
    std::call_once(once, [] {
        std::vector<std::string> names;
        if (const char* e = std::getenv("MFIC_HDF5_LIB")) names.push_back(e);
        names.insert(names.end(), {"libhdf5.so", "libhdf5_serial.so", "/opt/conda/lib/libhdf5.so"});
        for (const auto& n : names)
            if ((h.so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
        if (!h.so) {
            h.error = "HDF5 C library not found (set MFIC_HDF5_LIB to libhdf5.so)";
            return;
        }
        bool ok = sym(h.so, "H5open", h.open) && sym(h.so, "H5Fcreate", h.Fcreate) && sym(h.so, "H5Fclose", h.Fclose) &&
                  sym(h.so, "H5Gcreate2", h.Gcreate2) && sym(h.so, "H5Gclose", h.Gclose) &&
                  sym(h.so, "H5Screate", h.Screate) && sym(h.so, "H5Screate_simple", h.Screate_simple) &&
                  sym(h.so, "H5Sselect_hyperslab", h.Sselect_hyperslab) && sym(h.so, "H5Sclose", h.Sclose) &&
                  sym(h.so, "H5Dcreate2", h.Dcreate2) && sym(h.so, "H5Dwrite", h.Dwrite) &&
                  sym(h.so, "H5Dclose", h.Dclose) && sym(h.so, "H5Acreate2", h.Acreate2) &&
                  sym(h.so, "H5Awrite", h.Awrite) && sym(h.so, "H5Aclose", h.Aclose) && sym(h.so, "H5Tcopy", h.Tcopy) &&
                  sym(h.so, "H5Tset_size", h.Tset_size) && sym(h.so, "H5Tset_strpad", h.Tset_strpad) &&
                  sym(h.so, "H5Tclose", h.Tclose) && sym(h.so, "H5Eset_auto2", h.Eset_auto2) &&
                  sym(h.so, "H5Pcreate", h.Pcreate) && sym(h.so, "H5Pset_chunk", h.Pset_chunk) &&
                  sym(h.so, "H5Pset_fletcher32", h.Pset_fletcher32) && sym(h.so, "H5Pclose", h.Pclose);
        if (!ok || h.open() < 0) {
            h.error = "HDF5 C library lacks the 1.10 API";
            return;
        }
        // the native type ids are library globals, valid after H5open
        struct { const char* name; hid_t* dst; } types[] = {
            {"H5T_NATIVE_FLOAT_g", &h.FLOAT}, {"H5T_NATIVE_DOUBLE_g", &h.DOUBLE}, {"H5T_NATIVE_INT_g", &h.INT},
            {"H5T_NATIVE_UINT_g", &h.UINT},   {"H5T_NATIVE_ULLONG_g", &h.ULLONG}, {"H5T_C_S1_g", &h.C_S1},
            {"H5P_CLS_DATASET_CREATE_ID_g", &h.DATASET_CREATE}};
        for (auto& t : types) {
            const hid_t* p = static_cast<const hid_t*>(dlsym(h.so, t.name));
            if (!p) {
                h.error = std::string("HDF5 C library lacks ") + t.name;
                return;
            }
            *t.dst = *p;
        }
        h.Eset_auto2(0, nullptr, nullptr);  // errors come back as status codes
    });
    MFIC_CHECK(h.error.empty(), MFIC_ERR_UNSUPPORTED, h.error);
    return h;
}

#define H5CK(expr, what)                                                                   \
    do {                                                                                   \
        if ((expr) < 0) throw Error(MFIC_ERR_IO, std::string("HDF5: ") + (what) + " failed"); \
    } while (0)

// an open HDF5 object closed by its own close function
struct H5Obj {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    H5Obj(hid_t i, herr_t (*c)(hid_t), const char* what) : id(i), close(c) {
        if (id < 0) throw Error(MFIC_ERR_IO, std::string("HDF5: ") + what + " failed");
    }
    ~H5Obj() {
        if (id >= 0) close(id);
    }
    H5Obj(const H5Obj&) = delete;
    H5Obj& operator=(const H5Obj&) = delete;
};

// HDFWriteGroupAttribute with a std::vector (HDF_IO.hh:1084-1109): 1-D
static void attr_array(Hdf5& h, hid_t grp, const char* name, hid_t type, const void* data, size_t n) {
    const hsize_t dims = n;
    H5Obj sp(h.Screate_simple(1, &dims, nullptr), h.Sclose, "dataspace");
    H5Obj a(h.Acreate2(grp, name, type, sp.id, 0, 0), h.Aclose, name);
    H5CK(h.Awrite(a.id, type, data), name);
}
// the scalar form (HDF_IO.hh:1141-1166)
static void attr_scalar(Hdf5& h, hid_t grp, const char* name, hid_t type, const void* data) {
    H5Obj sp(h.Screate(0), h.Sclose, "dataspace");
    H5Obj a(h.Acreate2(grp, name, type, sp.id, 0, 0), h.Aclose, name);
    H5CK(h.Awrite(a.id, type, data), name);
}
// string attribute: fixed length, null padded (HDF_IO.hh:1229-1260)
static void attr_string(Hdf5& h, hid_t grp, const char* name, const std::string& s) {
    H5Obj t(h.Tcopy(h.C_S1), h.Tclose, "string type");
    H5CK(h.Tset_size(t.id, s.size()), "string size");
    H5CK(h.Tset_strpad(t.id, 1), "string padding");
    H5Obj sp(h.Screate(0), h.Sclose, "dataspace");
    H5Obj a(h.Acreate2(grp, name, t.id, sp.id, 0, 0), h.Aclose, name);
    H5CK(h.Awrite(a.id, t.id, s.c_str()), name);
}

// a dataset of the particle group, chunked with a Fletcher-32 checksum when
// `filtered` (HDFCreateEmptyDataset(Vector) with filter = true,
// HDF_IO.hh:888-963: 1 MB chunks of scalars, 256 KB rows of vectors; a chunk
// is cut to the dataset's length, which HDF5 requires of fixed-size data)
static hid_t create_dataset(Hdf5& h, hid_t grp, const char* name, hid_t type, size_t type_size, hid_t space,
                            bool vector, size_t n, bool filtered) {
    if (!filtered) return h.Dcreate2(grp, name, type, space, 0, 0, 0);
    H5Obj pl(h.Pcreate(h.DATASET_CREATE), h.Pclose, "property list");
    const hsize_t rows = std::max<hsize_t>(1, std::min<hsize_t>(n, (vector ? 256 * 1024 : 1024 * 1024) / type_size));
    const hsize_t chunk[2] = {rows, 3};
    H5CK(h.Pset_chunk(pl.id, vector ? 2 : 1, chunk), "chunk");
    H5CK(h.Pset_fletcher32(pl.id), "fletcher32");
    return h.Dcreate2(grp, name, type, space, 0, pl.id, 0);
}

// what one particle group of a file holds
struct GroupSpec {
    const char* group;      // "PartType1", ...
    bool f64;               // Coordinates / Velocities as 64-bit floats
    bool long_ids;          // uint64 ParticleIDs
    bool filtered;          // chunked + Fletcher-32 datasets
    std::vector<std::pair<const char*, double>> constants;  // [N] datasets of one value (Masses, ...)
    bool const_f64 = true;  // the constant datasets as 64-bit (else 32-bit) floats
};

// Coordinates, Velocities, ParticleIDs and the constant datasets of one
// species, streamed chunk by chunk from the source
static void write_particle_group(Hdf5& h, hid_t file, const ParticleSource& src, const GroupSpec& gs) {
    const size_t N = src.npart;
    H5Obj g(h.Gcreate2(file, gs.group, 0, 0, 0), h.Gclose, (std::string("group ") + gs.group).c_str());
    const hsize_t vdims[2] = {N, 3}, idims[1] = {N};
    H5Obj vsp(h.Screate_simple(2, vdims, nullptr), h.Sclose, "dataspace");
    H5Obj isp(h.Screate_simple(1, idims, nullptr), h.Sclose, "dataspace");
    const hid_t vtype = gs.f64 ? h.DOUBLE : h.FLOAT, idtype = gs.long_ids ? h.ULLONG : h.UINT;
    const size_t vsize = gs.f64 ? 8 : 4, idsize = gs.long_ids ? 8 : 4;
    H5Obj dpos(create_dataset(h, g.id, "Coordinates", vtype, vsize, vsp.id, true, N, gs.filtered), h.Dclose,
               "dataset Coordinates");
    H5Obj dvel(create_dataset(h, g.id, "Velocities", vtype, vsize, vsp.id, true, N, gs.filtered), h.Dclose,
               "dataset Velocities");
    H5Obj dids(create_dataset(h, g.id, "ParticleIDs", idtype, idsize, isp.id, false, N, gs.filtered), h.Dclose,
               "dataset ParticleIDs");
    std::vector<std::unique_ptr<H5Obj>> dconst;
    const hid_t ctype = gs.const_f64 ? h.DOUBLE : h.FLOAT;
    for (const auto& c : gs.constants)
        dconst.emplace_back(new H5Obj(create_dataset(h, g.id, c.first, ctype, gs.const_f64 ? 8 : 4, isp.id, false, N,
                                                     gs.filtered),
                                      h.Dclose, c.first));

    // chunk by chunk: [pos | vel | ids] per chunk
    const size_t chunk = std::min<size_t>(std::max<size_t>(N, 1), size_t(1) << 22);
    const size_t vbytes = 3 * vsize * chunk, chunk_bytes = 2 * vbytes + idsize * chunk;
    std::vector<double> cvals(gs.constants.empty() ? 0 : chunk);
    struct Buf {
        void* host = nullptr;
        void* dev = nullptr;
        hipEvent_t ready = nullptr;
    } buf[2];
    auto release = [&] {
        for (auto& b : buf) {
            if (b.ready) (void)hipEventDestroy(b.ready);
            if (b.host) {
                if (src.ctx) (void)hipHostFree(b.host);
                else std::free(b.host);
            }
            if (b.dev) (void)hipFree(b.dev);
        }
    };
    try {
        for (auto& b : buf) {
            if (src.ctx) {
                MFIC_HIP(hipHostMalloc(&b.host, chunk_bytes, hipHostMallocDefault));
                MFIC_HIP(hipMalloc(&b.dev, chunk_bytes));
                MFIC_HIP(hipEventCreateWithFlags(&b.ready, hipEventDisableTiming));
            } else {
                b.host = std::malloc(chunk_bytes);
                MFIC_CHECK(b.host, MFIC_ERR_ALLOC, "host staging buffer");
            }
        }
        const size_t nchunks = N ? (N + chunk - 1) / chunk : 0;
        // fill chunk c into buffer c % 2 (device: enqueue kernels and the
        // copies, record the event; host: convert in place)
        auto stage = [&](size_t c) {
            Buf& b = buf[c % 2];
            const size_t p0 = c * chunk, cnt = std::min(chunk, N - p0);
            char* hb = static_cast<char*>(b.host);
            if (src.ctx) {
                char* d = static_cast<char*>(b.dev);
                const unsigned bs = 256;
                const size_t n3 = 3 * cnt;
                const hipStream_t st = src.ctx->stream;
                if (gs.f64) {
                    // FP64 vectors leave as they are: straight device-to-host copies
                    MFIC_HIP(hipMemcpyAsync(hb, src.pos + 3 * p0, 24 * cnt, hipMemcpyDeviceToHost, st));
                    MFIC_HIP(hipMemcpyAsync(hb + vbytes, src.vel + 3 * p0, 24 * cnt, hipMemcpyDeviceToHost, st));
                } else {
                    hipLaunchKernelGGL(k_stage_f32, dim3((unsigned)((n3 + bs - 1) / bs)), dim3(bs), 0, st,
                                       src.pos + 3 * p0, reinterpret_cast<float*>(d), n3);
                    hipLaunchKernelGGL(k_stage_f32, dim3((unsigned)((n3 + bs - 1) / bs)), dim3(bs), 0, st,
                                       src.vel + 3 * p0, reinterpret_cast<float*>(d + vbytes), n3);
                    MFIC_LAUNCH_CHECK();
                    MFIC_HIP(hipMemcpyAsync(hb, d, 12 * cnt, hipMemcpyDeviceToHost, st));
                    MFIC_HIP(hipMemcpyAsync(hb + vbytes, d + vbytes, 12 * cnt, hipMemcpyDeviceToHost, st));
                }
                if (gs.long_ids)
                    hipLaunchKernelGGL(k_stage_ids<uint64_t>, dim3((unsigned)((cnt + bs - 1) / bs)), dim3(bs), 0, st,
                                       reinterpret_cast<uint64_t*>(d + 2 * vbytes), p0, cnt, src.ncell, src.overload,
                                       src.id_offset);
                else
                    hipLaunchKernelGGL(k_stage_ids<uint32_t>, dim3((unsigned)((cnt + bs - 1) / bs)), dim3(bs), 0, st,
                                       reinterpret_cast<uint32_t*>(d + 2 * vbytes), p0, cnt, src.ncell, src.overload,
                                       src.id_offset);
                MFIC_LAUNCH_CHECK();
                MFIC_HIP(hipMemcpyAsync(hb + 2 * vbytes, d + 2 * vbytes, idsize * cnt, hipMemcpyDeviceToHost, st));
                MFIC_HIP(hipEventRecord(b.ready, st));
            } else {
                for (size_t i = 0; i < 3 * cnt; ++i) {
                    if (gs.f64) {
                        reinterpret_cast<double*>(hb)[i] = src.pos[3 * p0 + i];
                        reinterpret_cast<double*>(hb + vbytes)[i] = src.vel[3 * p0 + i];
                    } else {
                        reinterpret_cast<float*>(hb)[i] = (float)src.pos[3 * p0 + i];
                        reinterpret_cast<float*>(hb + vbytes)[i] = (float)src.vel[3 * p0 + i];
                    }
                }
                char* ids = hb + 2 * vbytes;
                for (size_t i = 0; i < cnt; ++i) {
                    const size_t p = p0 + i, iload = p / src.ncell, cell = p - iload * src.ncell;
                    const uint64_t id = src.id_offset + (uint64_t)src.overload * cell + iload;
                    if (gs.long_ids) reinterpret_cast<uint64_t*>(ids)[i] = id;
                    else reinterpret_cast<uint32_t*>(ids)[i] = (uint32_t)id;
                }
            }
        };
        if (nchunks) stage(0);
        for (size_t c = 0; c < nchunks; ++c) {
            if (c + 1 < nchunks) stage(c + 1);  // device: overlaps this chunk's writes
            Buf& b = buf[c % 2];
            if (src.ctx) MFIC_HIP(hipEventSynchronize(b.ready));
            const size_t p0 = c * chunk, cnt = std::min(chunk, N - p0);
            const hsize_t vstart[2] = {p0, 0}, vcount[2] = {cnt, 3}, istart[1] = {p0}, icount[1] = {cnt};
            H5Obj vmem(h.Screate_simple(2, vcount, nullptr), h.Sclose, "memory dataspace");
            H5Obj imem(h.Screate_simple(1, icount, nullptr), h.Sclose, "memory dataspace");
            H5CK(h.Sselect_hyperslab(vsp.id, 0, vstart, nullptr, vcount, nullptr), "select");
            H5CK(h.Sselect_hyperslab(isp.id, 0, istart, nullptr, icount, nullptr), "select");
            char* hb = static_cast<char*>(b.host);
            H5CK(h.Dwrite(dpos.id, vtype, vmem.id, vsp.id, 0, hb), "write Coordinates");
            H5CK(h.Dwrite(dvel.id, vtype, vmem.id, vsp.id, 0, hb + vbytes), "write Velocities");
            H5CK(h.Dwrite(dids.id, idtype, imem.id, isp.id, 0, hb + 2 * vbytes), "write ParticleIDs");
            for (size_t k = 0; k < gs.constants.size(); ++k) {
                // the value as the container holds it (mass32_ or mass64_), converted by HDF5
                const double v = gs.const_f64 ? gs.constants[k].second : (double)(float)gs.constants[k].second;
                std::fill(cvals.begin(), cvals.begin() + cnt, v);
                H5CK(h.Dwrite(dconst[k]->id, h.DOUBLE, imem.id, isp.id, 0, cvals.data()), gs.constants[k].first);
            }
            // the next stage into this buffer must wait for these writes:
            // H5Dwrite has returned, so the host buffer is free again
        }
    } catch (...) {
        if (src.ctx) (void)hipStreamSynchronize(src.ctx->stream);
        release();
        throw;
    }
    if (src.ctx) MFIC_HIP(hipStreamSynchronize(src.ctx->stream));
    release();
}

// /ICs_parameters of both plugins (output_gadget_hdf5.cc:136-170,
// output_swift.cc:135-177); the git keys carry the library version
static void write_ics_parameters(Hdf5& h, hid_t file, const mfic_ic_config& cfg, bool tabulated) {
    H5Obj g(h.Gcreate2(file, "ICs_parameters", 0, 0, 0), h.Gclose, "group ICs_parameters");
    static const char* loads[] = {"sc", "bcc", "fcc"};
    const int order = cfg.lpt_order, fix = cfg.do_fixing != 0, inv = cfg.do_inversion != 0, zero = 0,
              bar = cfg.baryons != 0, L = (int)cfg.n;
    attr_string(h, g.id, "Code", "monofonic_amd (MI355X)");
    attr_string(h, g.id, "Git Revision", mfic_version());
    attr_string(h, g.id, "Git Tag", mfic_version());
    attr_string(h, g.id, "Git Branch", "none");
    attr_string(h, g.id, "Precision", cfg.prec == MFIC_F64 ? "double" : "single");
    attr_string(h, g.id, "Convolutions", "orszag");
    attr_string(h, g.id, "PLT", cfg.plt ? "yes" : "no");
    attr_scalar(h, g.id, "LPT Order", h.INT, &order);
    attr_string(h, g.id, "Particle Load", loads[cfg.lattice]);
    attr_string(h, g.id, "Transfer Function", tabulated ? "CAMB_file" : "eisenstein");
    attr_string(h, g.id, "Cosmology Parameter Set", "none");
    attr_string(h, g.id, "Random Generator", "NGENIC");
    attr_scalar(h, g.id, "Mode Fixing", h.INT, &fix);
    attr_scalar(h, g.id, "Mode inversion", h.INT, &inv);
    attr_scalar(h, g.id, "Baryons", h.INT, &bar);
    attr_scalar(h, g.id, "Baryons Relative Velocity", h.INT, &zero);
    attr_scalar(h, g.id, "Grid Resolution", h.INT, &L);
}

void gadget_write(const char* fname, const mfic_ic_config& cfg, const mfic_cosmo_params& cosmo,
                  const mfic_gadget_opts& opts, const ParticleSource& src, const ParticleSource* bsrc, bool tabulated) {
    MFIC_CHECK(fname && *fname, MFIC_ERR_ARG, "null file name");
    Hdf5& h = hdf5();
    mfic_cosmo_scalars sc;
    cosmo_compute(cosmo, &sc, nullptr);
    double lunit, vunit, munit;
    gadget_units(cfg.L, cosmo.zstart, &lunit, &vunit, &munit);
    const double astart = 1.0 / (1.0 + cosmo.zstart);
    const size_t N = src.npart;

    // unlink + HDFCreateFile: H5F_ACC_TRUNC replaces an existing file
    H5Obj file(h.Fcreate(fname, 0x0002u, 0, 0), h.Fclose, (std::string("create ") + fname).c_str());
    write_ics_parameters(h, file.id, cfg, tabulated);
    // species masses (ic_generator.cc:157-166, no neutrino exclusion)
    const bool bar = bsrc != nullptr;
    const double Omega_dm = bar ? cosmo.Omega_m - cosmo.Omega_b : cosmo.Omega_m;
    const size_t Nb = bar ? bsrc->npart : 0;
    // PartType1 (dark matter, get_species_idx) with Coordinates, Velocities
    // and ParticleIDs (output_gadget_hdf5.cc:268-318); with baryons every
    // particle carries its mass (set_masses, ic_generator.cc:747-778), so
    // both species add Masses in the plugin's precision and the MassTable
    // stays zero
    GroupSpec dm{"PartType1", false, opts.long_ids != 0, false, {}};
    if (bar) {
        dm.constants.push_back({"Masses", N ? Omega_dm * munit / (double)N : 0.0});
        dm.const_f64 = opts.double_precision != 0;
    }
    write_particle_group(h, file.id, src, dm);
    if (bar) {
        GroupSpec gb{"PartType0", false, opts.long_ids != 0, false, {{"Masses", Nb ? cosmo.Omega_b * munit / (double)Nb : 0.0}}};
        gb.const_f64 = opts.double_precision != 0;
        write_particle_group(h, file.id, *bsrc, gb);
    }
    {
        // Header (output_gadget_hdf5.cc:175-210)
        H5Obj g(h.Gcreate2(file.id, "Header", 0, 0, 0), h.Gclose, "group Header");
        double mass[6] = {0, 0, 0, 0, 0, 0};
        if (!bar) mass[1] = N ? Omega_dm * munit / (double)N : 0.0;
        if (opts.gadget2_compat) {
            unsigned np[6] = {(unsigned)Nb, (unsigned)N, 0, 0, 0, 0},
                     hw[6] = {(unsigned)((uint64_t)Nb >> 32), (unsigned)((uint64_t)N >> 32), 0, 0, 0, 0};
            attr_array(h, g.id, "NumPart_ThisFile", h.UINT, np, 6);
            attr_array(h, g.id, "NumPart_Total", h.UINT, np, 6);
            attr_array(h, g.id, "NumPart_Total_HighWord", h.UINT, hw, 6);
        } else {
            unsigned long long np[6] = {(unsigned long long)Nb, (unsigned long long)N, 0, 0, 0, 0};
            attr_array(h, g.id, "NumPart_ThisFile", h.ULLONG, np, 6);
            attr_array(h, g.id, "NumPart_Total", h.ULLONG, np, 6);
        }
        const double time = astart, redshift = 1.0 / astart - 1.0;
        const int izero = 0, one = 1;
        attr_array(h, g.id, "MassTable", h.DOUBLE, mass, 6);
        attr_array(h, g.id, "Time", h.DOUBLE, &time, 1);
        attr_array(h, g.id, "Redshift", h.DOUBLE, &redshift, 1);
        attr_array(h, g.id, "Flag_Sfr", h.INT, &izero, 1);
        attr_array(h, g.id, "Flag_Feedback", h.INT, &izero, 1);
        attr_array(h, g.id, "Flag_Cooling", h.INT, &izero, 1);
        attr_array(h, g.id, "NumFilesPerSnapshot", h.INT, &one, 1);
        attr_array(h, g.id, "BoxSize", h.DOUBLE, &lunit, 1);
        attr_array(h, g.id, "Omega0", h.DOUBLE, &cosmo.Omega_m, 1);
        attr_array(h, g.id, "OmegaLambda", h.DOUBLE, &sc.Omega_DE, 1);
        attr_array(h, g.id, "HubbleParam", h.DOUBLE, &sc.h, 1);
        attr_array(h, g.id, "Flag_StellarAge", h.INT, &izero, 1);
        attr_array(h, g.id, "Flag_Metals", h.INT, &izero, 1);
        attr_array(h, g.id, "Flag_Entropy_ICs", h.INT, &izero, 1);
    }
}

void swift_write(const char* fname, const mfic_ic_config& cfg, const mfic_cosmo_params& cosmo,
                 const mfic_swift_opts& opts, const ParticleSource& src, const ParticleSource* bsrc, bool tabulated) {
    MFIC_CHECK(fname && *fname, MFIC_ERR_ARG, "null file name");
    Hdf5& h = hdf5();
    mfic_cosmo_scalars sc;
    cosmo_compute(cosmo, &sc, nullptr);
    double lunit, vunit, munit;
    swift_units(cfg.L, cosmo, &lunit, &vunit, &munit);
    const size_t N = src.npart;

    H5Obj file(h.Fcreate(fname, 0x0002u, 0, 0), h.Fclose, (std::string("create ") + fname).c_str());
    {
        // the physical constants SWIFT assumes internally (output_swift.cc:124-130)
        H5Obj g(h.Gcreate2(file.id, "Units", 0, 0, 0), h.Gclose, "group Units");
        const double um = 1.98841e43, ul = 3.08567758149e24, ut = 3.08567758149e19, one = 1.0;
        attr_scalar(h, g.id, "Unit mass in cgs (U_M)", h.DOUBLE, &um);
        attr_scalar(h, g.id, "Unit length in cgs (U_L)", h.DOUBLE, &ul);
        attr_scalar(h, g.id, "Unit time in cgs (U_t)", h.DOUBLE, &ut);
        attr_scalar(h, g.id, "Unit current in cgs (U_I)", h.DOUBLE, &one);
        attr_scalar(h, g.id, "Unit temperature in cgs (U_T)", h.DOUBLE, &one);
    }
    write_ics_parameters(h, file.id, cfg, tabulated);
    // PartType1 with per-particle Masses (output_swift.cc:250-377): every
    // particle has Omega_s munit / N_s, whether the container holds the
    // masses (with baryons) or the plugin fills them in
    const bool bar = bsrc != nullptr;
    const double Omega_dm = bar ? cosmo.Omega_m - cosmo.Omega_b : cosmo.Omega_m;
    const size_t Nb = bar ? bsrc->npart : 0;
    const double pmass = N ? Omega_dm * munit / (double)N : 0.0;
    write_particle_group(h, file.id, src, GroupSpec{"PartType1", true, opts.long_ids != 0, true, {{"Masses", pmass}}});
    if (bar) {
        // PartType0: gas with the initial internal energy and the mean
        // inter-particle separation as smoothing length (output_swift.cc:
        // 86-110, 297-302, 360-366)
        const double astart = 1.0 / (1.0 + cosmo.zstart), hh = sc.h;
        const double gamma = opts.gamma > 0 ? opts.gamma : 5.0 / 3.0, YHe = opts.YHe > 0 ? opts.YHe : 0.245421;
        const double Tcmb0 = cosmo.Tcmb;
        const double npol = (std::fabs(1.0 - gamma) > 1e-7) ? 1.0 / (gamma - 1.) : 1.0;
        const double unitv = 1e5;
        const double adec = 1.0 / (160. * std::pow(cosmo.Omega_b * hh * hh / 0.022, 2.0 / 5.0));
        const double Tini = astart < adec ? Tcmb0 / astart : Tcmb0 / astart / astart * adec;
        const double mu = (Tini > 1.e4) ? 4.0 / (8. - 5. * YHe) : 4.0 / (1. + 3. * (1. - YHe));
        const double ceint = 1.3806e-16 / 1.6726e-24 * Tini * npol / mu / unitv / unitv;
        const double hsml = cfg.L / hh / (double)cfg.n;
        write_particle_group(h, file.id, *bsrc,
                             GroupSpec{"PartType0", true, opts.long_ids != 0, true,
                                       {{"Masses", Nb ? cosmo.Omega_b * munit / (double)Nb : 0.0},
                                        {"InternalEnergy", ceint},
                                        {"SmoothingLength", hsml}}});
    }
    {
        // Header, written by the plugin's destructor (output_swift.cc:183-200);
        // its MassTable stays zero (masses are per particle)
        H5Obj g(h.Gcreate2(file.id, "Header", 0, 0, 0), h.Gclose, "group Header");
        const int dim = 3, izero = 0, one = 1;
        const double box = cfg.L / sc.h, time = 1.0 / (1.0 + cosmo.zstart), redshift = cosmo.zstart;
        unsigned tot[7] = {(unsigned)Nb, (unsigned)N, 0, 0, 0, 0, 0},
                 hw[7] = {(unsigned)((uint64_t)Nb >> 32), (unsigned)((uint64_t)N >> 32), 0, 0, 0, 0, 0};
        unsigned long long thisfile[7] = {(unsigned long long)Nb, (unsigned long long)N, 0, 0, 0, 0, 0};
        const double mass[7] = {0, 0, 0, 0, 0, 0, 0};
        attr_scalar(h, g.id, "Dimension", h.INT, &dim);
        attr_scalar(h, g.id, "BoxSize", h.DOUBLE, &box);
        attr_array(h, g.id, "NumPart_Total", h.UINT, tot, 7);
        attr_array(h, g.id, "NumPart_Total_HighWord", h.UINT, hw, 7);
        attr_array(h, g.id, "NumPart_ThisFile", h.ULLONG, thisfile, 7);
        attr_array(h, g.id, "MassTable", h.DOUBLE, mass, 7);
        attr_array(h, g.id, "Time", h.DOUBLE, &time, 1);
        attr_array(h, g.id, "Redshift", h.DOUBLE, &redshift, 1);
        attr_array(h, g.id, "Flag_Entropy_ICs", h.INT, &izero, 1);
        attr_array(h, g.id, "NumFilesPerSnapshot", h.INT, &one, 1);
    }
}

void swift_units(double boxlength, const mfic_cosmo_params& cosmo, double* lunit, double* vunit, double* munit) {
    // output_swift.cc:63-69: Mpc (no h), km/s, 1e10 Msun with SWIFT's own
    // rho_crit in h^2 1e10 Msun / Mpc^3
    const double h = cosmo.H0 / 100.0, rhoc = 27.7536609198;
    *lunit = boxlength / h;
    *vunit = boxlength;
    *munit = rhoc * std::pow(boxlength, 3) / h;
}

void gadget_units(double boxlength, double zstart, double* lunit, double* vunit, double* munit) {
    // output_gadget_hdf5.cc:79-85: rho_crit in h^2 1e10 Msun / Mpc^3
    const double astart = 1.0 / (1.0 + zstart);
    const double rhoc = 27.7519737;
    *lunit = boxlength;
    *vunit = boxlength / std::sqrt(astart);
    *munit = rhoc * std::pow(boxlength, 3);
}

}  // namespace mfic
