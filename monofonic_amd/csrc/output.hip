// Particle output: the Gadget-HDF5 writer of the gadget_hdf5 /
// gadget_hdf5_double output plugins (output_gadget_hdf5.cc:36-318) fed
// straight from the device particle arrays of a run.
//
// File layout and datatypes follow the plugin and HDF_IO.hh:
//   /ICs_parameters  scalar attributes describing the run
//                    (output_gadget_hdf5.cc:140-170, HDF_IO.hh:1141-1262)
//   /PartType1/Coordinates, /Velocities  [N][3], stored as 32-bit floats
//                    whatever the plugin's precision: HDF_IO.hh:850-885
//                    creates vector datasets with H5T_NATIVE_FLOAT
//   /PartType1/ParticleIDs  [N] uint32 or (UseLongids) uint64
//                    (HDF_IO.hh:627-652), id = 2^lattice * cell + sub-lattice
//                    (particle_generator.hh:229-251)
//   /Header          1-element attribute arrays (output_gadget_hdf5.cc:175-210)
// The device converts each chunk of FP64 [p][3] values to the file's 32-bit
// floats and generates the IDs, so PCIe carries only the bytes written; two
// pinned staging buffers let the copy of chunk i+1 run while HDF5 writes
// chunk i.
//
// The HDF5 C library (1.10 API) is loaded at run time with dlopen, so the
// library builds and runs without it until a file is written: the path in
// MFIC_HDF5_LIB, then libhdf5.so on the loader path, then the conda copy.
#include <dlfcn.h>

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
// has id overload * cell + iload (particle_generator.hh:229-251, IDoffset 0)
template <typename I>
__global__ void k_stage_ids(I* __restrict__ dst, size_t p0, size_t cnt, size_t ncell, unsigned overload) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const size_t p = p0 + i;
    const size_t iload = p / ncell, cell = p - iload * ncell;
    dst[i] = (I)(overload * cell + iload);
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
    hid_t FLOAT = -1, DOUBLE = -1, INT = -1, UINT = -1, ULLONG = -1, C_S1 = -1;
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
                  sym(h.so, "H5Tclose", h.Tclose) && sym(h.so, "H5Eset_auto2", h.Eset_auto2);
        if (!ok || h.open() < 0) {
            h.error = "HDF5 C library lacks the 1.10 API";
            return;
        }
        // the native type ids are library globals, valid after H5open
        struct { const char* name; hid_t* dst; } types[] = {
            {"H5T_NATIVE_FLOAT_g", &h.FLOAT}, {"H5T_NATIVE_DOUBLE_g", &h.DOUBLE}, {"H5T_NATIVE_INT_g", &h.INT},
            {"H5T_NATIVE_UINT_g", &h.UINT},   {"H5T_NATIVE_ULLONG_g", &h.ULLONG}, {"H5T_C_S1_g", &h.C_S1}};
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

void gadget_write(const char* fname, const mfic_ic_config& cfg, const mfic_cosmo_params& cosmo,
                  const mfic_gadget_opts& opts, const ParticleSource& src, bool tabulated) {
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
    {
        // ICs_parameters (output_gadget_hdf5.cc:140-170)
        H5Obj g(h.Gcreate2(file.id, "ICs_parameters", 0, 0, 0), h.Gclose, "group ICs_parameters");
        static const char* loads[] = {"sc", "bcc", "fcc"};
        const int order = cfg.lpt_order, fix = cfg.do_fixing != 0, inv = cfg.do_inversion != 0, zero = 0,
                  L = (int)cfg.n;
        attr_string(h, g.id, "Code", "monofonic_amd (MI355X)");
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
        attr_scalar(h, g.id, "Baryons", h.INT, &zero);
        attr_scalar(h, g.id, "Baryons Relative Velocity", h.INT, &zero);
        attr_scalar(h, g.id, "Grid Resolution", h.INT, &L);
    }
    {
        // PartType1 (dark matter, get_species_idx) with Coordinates,
        // Velocities and ParticleIDs (output_gadget_hdf5.cc:268-318)
        H5Obj g(h.Gcreate2(file.id, "PartType1", 0, 0, 0), h.Gclose, "group PartType1");
        const hsize_t vdims[2] = {N, 3}, idims[1] = {N};
        H5Obj vsp(h.Screate_simple(2, vdims, nullptr), h.Sclose, "dataspace");
        H5Obj isp(h.Screate_simple(1, idims, nullptr), h.Sclose, "dataspace");
        const hid_t idtype = opts.long_ids ? h.ULLONG : h.UINT;
        const size_t idsize = opts.long_ids ? 8 : 4;
        H5Obj dpos(h.Dcreate2(g.id, "Coordinates", h.FLOAT, vsp.id, 0, 0, 0), h.Dclose, "dataset Coordinates");
        H5Obj dvel(h.Dcreate2(g.id, "Velocities", h.FLOAT, vsp.id, 0, 0, 0), h.Dclose, "dataset Velocities");
        H5Obj dids(h.Dcreate2(g.id, "ParticleIDs", idtype, isp.id, 0, 0, 0), h.Dclose, "dataset ParticleIDs");

        // chunk by chunk: [pos f32 | vel f32 | ids] per chunk
        const size_t chunk = std::min<size_t>(N, size_t(1) << 22);
        const size_t chunk_bytes = chunk * (3 * 4 * 2 + idsize);
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
            // copy, record the event; host: convert in place)
            auto stage = [&](size_t c) {
                Buf& b = buf[c % 2];
                const size_t p0 = c * chunk, cnt = std::min(chunk, N - p0);
                if (src.ctx) {
                    char* d = static_cast<char*>(b.dev);
                    const unsigned bs = 256;
                    const size_t n3 = 3 * cnt;
                    hipLaunchKernelGGL(k_stage_f32, dim3((unsigned)((n3 + bs - 1) / bs)), dim3(bs), 0, src.ctx->stream,
                                       src.pos + 3 * p0, reinterpret_cast<float*>(d), n3);
                    hipLaunchKernelGGL(k_stage_f32, dim3((unsigned)((n3 + bs - 1) / bs)), dim3(bs), 0, src.ctx->stream,
                                       src.vel + 3 * p0, reinterpret_cast<float*>(d + 12 * chunk), n3);
                    if (opts.long_ids)
                        hipLaunchKernelGGL(k_stage_ids<uint64_t>, dim3((unsigned)((cnt + bs - 1) / bs)), dim3(bs), 0,
                                           src.ctx->stream, reinterpret_cast<uint64_t*>(d + 24 * chunk), p0, cnt,
                                           src.ncell, src.overload);
                    else
                        hipLaunchKernelGGL(k_stage_ids<uint32_t>, dim3((unsigned)((cnt + bs - 1) / bs)), dim3(bs), 0,
                                           src.ctx->stream, reinterpret_cast<uint32_t*>(d + 24 * chunk), p0, cnt,
                                           src.ncell, src.overload);
                    MFIC_LAUNCH_CHECK();
                    MFIC_HIP(hipMemcpyAsync(b.host, b.dev, 12 * cnt, hipMemcpyDeviceToHost, src.ctx->stream));
                    MFIC_HIP(hipMemcpyAsync(static_cast<char*>(b.host) + 12 * chunk, d + 12 * chunk, 12 * cnt,
                                            hipMemcpyDeviceToHost, src.ctx->stream));
                    MFIC_HIP(hipMemcpyAsync(static_cast<char*>(b.host) + 24 * chunk, d + 24 * chunk, idsize * cnt,
                                            hipMemcpyDeviceToHost, src.ctx->stream));
                    MFIC_HIP(hipEventRecord(b.ready, src.ctx->stream));
                } else {
                    float* hp = static_cast<float*>(b.host);
                    float* hv = reinterpret_cast<float*>(static_cast<char*>(b.host) + 12 * chunk);
                    for (size_t i = 0; i < 3 * cnt; ++i) {
                        hp[i] = (float)src.pos[3 * p0 + i];
                        hv[i] = (float)src.vel[3 * p0 + i];
                    }
                    char* ids = static_cast<char*>(b.host) + 24 * chunk;
                    for (size_t i = 0; i < cnt; ++i) {
                        const size_t p = p0 + i, iload = p / src.ncell, cell = p - iload * src.ncell;
                        const uint64_t id = (uint64_t)src.overload * cell + iload;
                        if (opts.long_ids) reinterpret_cast<uint64_t*>(ids)[i] = id;
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
                H5CK(h.Dwrite(dpos.id, h.FLOAT, vmem.id, vsp.id, 0, b.host), "write Coordinates");
                H5CK(h.Dwrite(dvel.id, h.FLOAT, vmem.id, vsp.id, 0, static_cast<char*>(b.host) + 12 * chunk),
                     "write Velocities");
                H5CK(h.Dwrite(dids.id, idtype, imem.id, isp.id, 0, static_cast<char*>(b.host) + 24 * chunk),
                     "write ParticleIDs");
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
    {
        // Header (output_gadget_hdf5.cc:175-210)
        H5Obj g(h.Gcreate2(file.id, "Header", 0, 0, 0), h.Gclose, "group Header");
        const double Omega_dm = cosmo.Omega_m;  // no baryons, no neutrino exclusion (ic_generator.cc:156-195)
        double mass[6] = {0, 0, 0, 0, 0, 0};
        mass[1] = N ? Omega_dm * munit / (double)N : 0.0;
        if (opts.gadget2_compat) {
            unsigned np[6] = {0, (unsigned)N, 0, 0, 0, 0}, hw[6] = {0, (unsigned)((uint64_t)N >> 32), 0, 0, 0, 0};
            attr_array(h, g.id, "NumPart_ThisFile", h.UINT, np, 6);
            attr_array(h, g.id, "NumPart_Total", h.UINT, np, 6);
            attr_array(h, g.id, "NumPart_Total_HighWord", h.UINT, hw, 6);
        } else {
            unsigned long long np[6] = {0, (unsigned long long)N, 0, 0, 0, 0};
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

void gadget_units(double boxlength, double zstart, double* lunit, double* vunit, double* munit) {
    // output_gadget_hdf5.cc:79-85: rho_crit in h^2 1e10 Msun / Mpc^3
    const double astart = 1.0 / (1.0 + zstart);
    const double rhoc = 27.7519737;
    *lunit = boxlength;
    *vunit = boxlength / std::sqrt(astart);
    *munit = rhoc * std::pow(boxlength, 3);
}

}  // namespace mfic
