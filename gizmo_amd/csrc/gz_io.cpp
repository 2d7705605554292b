// gz_io.cpp -- GIZMO HDF5 ICs and snapshots of the gas fields (include/gizmo_io.h), HDF5 C API.
//
// Layout as the reference writes and reads it: header attributes (src/file_io/io.cc:5324-5440,
// src/file_io/read_ic.cc:2204-2260) and the PartType0 dataset names of get_dataset_name
// (src/file_io/io.cc:4374-4800) with the legacy kernel-radius alias (io.cc:4848-4866).
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gizmo_io.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// closes an HDF5 handle on scope exit
struct H {
    hid_t id;
    herr_t (*close)(hid_t);
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H() { if (id >= 0) close(id); }
    operator hid_t() const { return id; }
};

// silence HDF5's own error printing for the probes of optional objects
struct Quiet {
    H5E_auto2_t f;
    void *d;
    Quiet() { H5Eget_auto2(H5E_DEFAULT, &f, &d); H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, f, d); }
};

bool has_attr(hid_t loc, const char *name) { return H5Aexists(loc, name) > 0; }
bool has_link(hid_t loc, const char *name) { return H5Lexists(loc, name, H5P_DEFAULT) > 0; }

template <typename T>
int read_attr(hid_t loc, const char *name, hid_t mem_type, T *out, int count)
{
    H a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    if (a < 0) return fail(GZ_E_STATE, "cannot open header attribute %s", name);
    H sp(H5Aget_space(a), H5Sclose);
    const hssize_t np = H5Sget_simple_extent_npoints(sp);
    if (np != count) return fail(GZ_E_STATE, "header attribute %s has %lld values, expected %d", name, (long long)np, count);
    if (H5Aread(a, mem_type, out) < 0) return fail(GZ_E_STATE, "cannot read header attribute %s", name);
    return 0;
}

int write_attr(hid_t loc, const char *name, hid_t file_type, hid_t mem_type, const void *v, int count)
{
    hsize_t dims[1] = {(hsize_t)count};
    H sp(count == 1 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, dims, nullptr), H5Sclose);
    H a(H5Acreate2(loc, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a < 0 || H5Awrite(a, mem_type, v) < 0) return fail(GZ_E_STATE, "cannot write header attribute %s", name);
    return 0;
}

int read_header(hid_t file, gz_snap_header *h)
{
    memset(h, 0, sizeof *h);
    Quiet q;
    H hdr(H5Gopen2(file, "Header", H5P_DEFAULT), H5Gclose);
    if (hdr < 0) return fail(GZ_E_STATE, "no Header group");
    long long npart[6];
    int rc = read_attr(hdr, "NumPart_ThisFile", H5T_NATIVE_LLONG, npart, 6);
    if (rc) return rc;
    double mt[6];
    if ((rc = read_attr(hdr, "MassTable", H5T_NATIVE_DOUBLE, mt, 6))) return rc;
    if ((rc = read_attr(hdr, "Time", H5T_NATIVE_DOUBLE, &h->time, 1))) return rc;
    h->num_gas = npart[0];
    h->mass_table_gas = mt[0];
    if (has_attr(hdr, "Redshift") && (rc = read_attr(hdr, "Redshift", H5T_NATIVE_DOUBLE, &h->redshift, 1))) return rc;
    if (has_attr(hdr, "BoxSize") && (rc = read_attr(hdr, "BoxSize", H5T_NATIVE_DOUBLE, &h->box_size, 1))) return rc;
    h->num_files = 1;
    if (has_attr(hdr, "NumFilesPerSnapshot") && (rc = read_attr(hdr, "NumFilesPerSnapshot", H5T_NATIVE_INT, &h->num_files, 1))) return rc;
    // ICs often lack it (read_ic.cc:2188): single precision is the reference's default then
    if (has_attr(hdr, "Flag_DoublePrecision")) {
        if ((rc = read_attr(hdr, "Flag_DoublePrecision", H5T_NATIVE_INT, &h->double_precision, 1))) return rc;
    }
    return 0;
}

// a PartType0 dataset of n rows x width columns into out (converted to FP64 or int64)
int read_dataset(hid_t grp, const char *name, long n, int width, hid_t mem_type, void *out)
{
    H d(H5Dopen2(grp, name, H5P_DEFAULT), H5Dclose);
    if (d < 0) return fail(GZ_E_STATE, "cannot open PartType0/%s", name);
    H sp(H5Dget_space(d), H5Sclose);
    hsize_t dims[2] = {0, 0};
    const int nd = H5Sget_simple_extent_ndims(sp);
    H5Sget_simple_extent_dims(sp, dims, nullptr);
    const bool shape_ok = (width == 1 && nd == 1 && (long)dims[0] == n) || (width > 1 && nd == 2 && (long)dims[0] == n && (int)dims[1] == width);
    if (!shape_ok) return fail(GZ_E_STATE, "PartType0/%s has shape (%llu, %llu), expected (%ld, %d)", name, (unsigned long long)dims[0],
                               nd > 1 ? (unsigned long long)dims[1] : 1ull, n, width);
    if (H5Dread(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) return fail(GZ_E_STATE, "cannot read PartType0/%s", name);
    return 0;
}

int write_dataset(hid_t grp, const char *name, long n, int width, bool fp64, const double *v)
{
    hsize_t dims[2] = {(hsize_t)n, (hsize_t)width};
    H sp(H5Screate_simple(width == 1 ? 1 : 2, dims, nullptr), H5Sclose);
    H d(H5Dcreate2(grp, name, fp64 ? H5T_IEEE_F64LE : H5T_IEEE_F32LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (d < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
        return fail(GZ_E_STATE, "cannot write PartType0/%s", name);
    return 0;
}

}  // namespace

extern "C" {

const char *gz_io_last_error(void) { return g_err.c_str(); }

int gz_ic_header(const char *path, gz_snap_header *hdr)
{
    if (!path || !hdr) return fail(GZ_E_ARG, "null argument");
    Quiet q;
    H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file < 0) return fail(GZ_E_ARG, "cannot open %s as HDF5", path);
    return read_header(file, hdr);
}

int gz_ic_read(const char *path, long n, gz_fields *f, long long *ids, int *has_h)
{
    if (!path || !f) return fail(GZ_E_ARG, "null argument");
    Quiet q;
    H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file < 0) return fail(GZ_E_ARG, "cannot open %s as HDF5", path);
    gz_snap_header hdr;
    int rc = read_header(file, &hdr);
    if (rc) return rc;
    if (hdr.num_files != 1) return fail(GZ_E_STATE, "%s is one file of %d; only single-file snapshots are read", path, hdr.num_files);
    if (hdr.num_gas != n) return fail(GZ_E_ARG, "%s holds %lld gas particles, %ld requested", path, hdr.num_gas, n);
    if (n == 0) return 0;
    H grp(H5Gopen2(file, "PartType0", H5P_DEFAULT), H5Gclose);
    if (grp < 0) return fail(GZ_E_STATE, "no PartType0 group");
    if (f->pos && (rc = read_dataset(grp, "Coordinates", n, 3, H5T_NATIVE_DOUBLE, f->pos))) return rc;
    if (f->vel && (rc = read_dataset(grp, "Velocities", n, 3, H5T_NATIVE_DOUBLE, f->vel))) return rc;
    if (f->u && (rc = read_dataset(grp, "InternalEnergy", n, 1, H5T_NATIVE_DOUBLE, f->u))) return rc;
    if (f->mass) {
        if (has_link(grp, "Masses")) {
            if ((rc = read_dataset(grp, "Masses", n, 1, H5T_NATIVE_DOUBLE, f->mass))) return rc;
        } else if (hdr.mass_table_gas > 0) {
            for (long k = 0; k < n; k++) f->mass[k] = hdr.mass_table_gas;
        } else {
            return fail(GZ_E_STATE, "no Masses dataset and MassTable[0] = 0");
        }
    }
    if (f->rho && has_link(grp, "Density") && (rc = read_dataset(grp, "Density", n, 1, H5T_NATIVE_DOUBLE, f->rho))) return rc;
    int hh = 0;
    if (f->h) {
        const char *name = has_link(grp, "KernelMaxRadius") ? "KernelMaxRadius" : (has_link(grp, "SmoothingLength") ? "SmoothingLength" : nullptr);
        if (name) {
            if ((rc = read_dataset(grp, name, n, 1, H5T_NATIVE_DOUBLE, f->h))) return rc;
            hh = 1;
        }
    }
    if (has_h) *has_h = hh;
    if (ids) {
        if (has_link(grp, "ParticleIDs")) {
            if ((rc = read_dataset(grp, "ParticleIDs", n, 1, H5T_NATIVE_LLONG, ids))) return rc;
        } else {
            for (long k = 0; k < n; k++) ids[k] = k + 1;
        }
    }
    return 0;
}

int gz_snapshot_write(const char *path, long n, const gz_fields *f, const long long *ids, const gz_snap_header *hdr)
{
    if (!path || !f || !hdr || n < 0) return fail(GZ_E_ARG, "null argument or negative count");
    if (!f->pos) return fail(GZ_E_ARG, "a snapshot needs positions");
    Quiet q;
    H file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file < 0) return fail(GZ_E_ARG, "cannot create %s", path);
    {
        H g(H5Gcreate2(file, "Header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (g < 0) return fail(GZ_E_STATE, "cannot create Header");
        const long long np[6] = {n, 0, 0, 0, 0, 0};
        const unsigned int nt[6] = {(unsigned int)(n & 0xffffffffLL), 0, 0, 0, 0, 0};
        const unsigned int nh[6] = {(unsigned int)((long long)n >> 32), 0, 0, 0, 0, 0};
        // every gas particle of one mass: MassTable[0] and no Masses dataset
        const double mt[6] = {f->mass ? 0.0 : hdr->mass_table_gas, 0, 0, 0, 0, 0};
        const int one = 1, zero = 0, dp = hdr->double_precision ? 1 : 0;
        const double hub = 1.0;
        int rc = 0;
        if ((rc = write_attr(g, "NumPart_ThisFile", H5T_NATIVE_LLONG, H5T_NATIVE_LLONG, np, 6)) ||
            (rc = write_attr(g, "NumPart_Total", H5T_NATIVE_UINT, H5T_NATIVE_UINT, nt, 6)) ||
            (rc = write_attr(g, "NumPart_Total_HighWord", H5T_NATIVE_UINT, H5T_NATIVE_UINT, nh, 6)) ||
            (rc = write_attr(g, "MassTable", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, mt, 6)) ||
            (rc = write_attr(g, "Time", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &hdr->time, 1)) ||
            (rc = write_attr(g, "Redshift", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &hdr->redshift, 1)) ||
            (rc = write_attr(g, "BoxSize", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &hdr->box_size, 1)) ||
            (rc = write_attr(g, "NumFilesPerSnapshot", H5T_NATIVE_INT, H5T_NATIVE_INT, &one, 1)) ||
            (rc = write_attr(g, "ComovingIntegrationOn", H5T_NATIVE_INT, H5T_NATIVE_INT, &zero, 1)) ||
            (rc = write_attr(g, "HubbleParam", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &hub, 1)) ||
            (rc = write_attr(g, "Flag_Sfr", H5T_NATIVE_INT, H5T_NATIVE_INT, &zero, 1)) ||
            (rc = write_attr(g, "Flag_Cooling", H5T_NATIVE_INT, H5T_NATIVE_INT, &zero, 1)) ||
            (rc = write_attr(g, "Flag_StellarAge", H5T_NATIVE_INT, H5T_NATIVE_INT, &zero, 1)) ||
            (rc = write_attr(g, "Flag_Metals", H5T_NATIVE_INT, H5T_NATIVE_INT, &zero, 1)) ||
            (rc = write_attr(g, "Flag_Feedback", H5T_NATIVE_INT, H5T_NATIVE_INT, &zero, 1)) ||
            (rc = write_attr(g, "Flag_DoublePrecision", H5T_NATIVE_INT, H5T_NATIVE_INT, &dp, 1)))
            return rc;
    }
    H grp(H5Gcreate2(file, "PartType0", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (grp < 0) return fail(GZ_E_STATE, "cannot create PartType0");
    const bool fp64 = hdr->double_precision != 0;
    struct Item { const char *name; const double *v; int width; };
    const Item items[] = {
        {"Coordinates", f->pos, 3},
        {"Velocities", f->vel_kick ? f->vel_kick : f->vel, 3},
        {"Masses", f->mass, 1},
        {"InternalEnergy", f->u_kick ? f->u_kick : f->u, 1},
        {"Density", f->rho, 1},
        {hdr->legacy_names ? "SmoothingLength" : "KernelMaxRadius", f->h, 1},
        {"Pressure", f->press, 1},
        {"HydroAcceleration", f->accel, 3},
        {"RateOfChangeOfInternalEnergy", f->dtu, 1},
        {"VelocityDivergence", f->divvel, 1},
        {"DensityGradient", f->grad_rho, 3},
        {"VelocityGradient", f->grad_v, 9},
    };
    for (const Item &it : items) {
        if (!it.v || n == 0) continue;
        int rc = write_dataset(grp, it.name, n, it.width, fp64, it.v);
        if (rc) return rc;
    }
    if (n > 0) {
        std::vector<long long> own;
        const long long *idv = ids;
        if (!idv) {
            own.resize(n);
            for (long k = 0; k < n; k++) own[k] = k + 1;
            idv = own.data();
        }
        hsize_t dims[1] = {(hsize_t)n};
        H sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
        H d(H5Dcreate2(grp, "ParticleIDs", H5T_STD_U64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        if (d < 0 || H5Dwrite(d, H5T_NATIVE_LLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, idv) < 0)
            return fail(GZ_E_STATE, "cannot write PartType0/ParticleIDs");
    }
    return 0;
}

}  // extern "C"
