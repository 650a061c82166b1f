// H5Part phase-space files of OPAL-T (include/opalsc_h5.h) on the HDF5 C API.
// Host code only, built into its own library (libopalsc_h5.so) so that the
// space-charge library carries no HDF5 dependency.
//
// The layout is H5hut's H5Part layer as H5PartWrapperForPT drives it
// (src/Structure/H5PartWrapperForPT.cpp): attributes on the root group, one
// group "Step#k" per dump holding the step attributes and one 1D dataset per
// particle quantity.  Strings are stored as H5hut stores them: a scalar
// dataspace of a fixed-length C string whose size counts the terminating NUL.
#include "../../include/opalsc_h5.h"

#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// Closes an HDF5 handle of one kind when the scope ends.
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    bool ok() const { return id >= 0; }
};

// HDF5's own error stack printing is off: failures come back as codes.
struct Quiet {
    H5E_auto2_t f = nullptr;
    void *d = nullptr;
    Quiet() {
        H5Eget_auto2(H5E_DEFAULT, &f, &d);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, f, d); }
};

bool write_string_attr(hid_t obj, const char *name, const char *value) {
    Hid type(H5Tcopy(H5T_C_S1), H5Tclose);
    if (!type.ok() || H5Tset_size(type.id, std::strlen(value) + 1) < 0) return false;
    Hid space(H5Screate(H5S_SCALAR), H5Sclose);
    Hid a(H5Acreate2(obj, name, type.id, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a.id, type.id, value) >= 0;
}

bool write_attr(hid_t obj, const char *name, hid_t type, const void *v, hsize_t n) {
    Hid space(H5Screate_simple(1, &n, nullptr), H5Sclose);
    Hid a(H5Acreate2(obj, name, type, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a.id, type, v) >= 0;
}
bool write_f64(hid_t obj, const char *name, const double *v, hsize_t n) {
    return write_attr(obj, name, H5T_NATIVE_DOUBLE, v, n);
}
bool write_i64(hid_t obj, const char *name, int64_t v) { return write_attr(obj, name, H5T_NATIVE_INT64, &v, 1); }

// Reads n values of a numeric attribute (converted to the memory type);
// false if it is absent or has another length.
bool read_attr(hid_t obj, const char *name, hid_t type, void *v, hsize_t n) {
    if (H5Aexists(obj, name) <= 0) return false;
    Hid a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return false;
    Hid space(H5Aget_space(a.id), H5Sclose);
    if (H5Sget_simple_extent_npoints(space.id) != (hssize_t)n) return false;
    return H5Aread(a.id, type, v) >= 0;
}

bool write_dataset(hid_t grp, const char *name, hid_t type, const void *data, hsize_t n) {
    Hid space(H5Screate_simple(1, &n, nullptr), H5Sclose);
    Hid d(H5Dcreate2(grp, name, type, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && H5Dwrite(d.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

// Elements [first, first + count) of a 1D dataset, converted to the memory type.
bool read_slab(hid_t grp, const char *name, hid_t type, hsize_t first, hsize_t count, void *out) {
    Hid d(H5Dopen2(grp, name, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return false;
    Hid fs(H5Dget_space(d.id), H5Sclose);
    if (H5Sselect_hyperslab(fs.id, H5S_SELECT_SET, &first, nullptr, &count, nullptr) < 0) return false;
    Hid ms(H5Screate_simple(1, &count, nullptr), H5Sclose);
    return H5Dread(d.id, type, ms.id, fs.id, H5P_DEFAULT, out) >= 0;
}

std::string step_name(int64_t k) { return "Step#" + std::to_string((long long)k); }

// Number of Step# groups directly under the root (H5GetNumSteps).
int64_t count_steps(hid_t file) {
    struct Ctx {
        int64_t n = 0;
    } ctx;
    auto cb = [](hid_t, const char *name, const H5L_info_t *, void *op) -> herr_t {
        if (std::strncmp(name, "Step#", 5) == 0) ++((Ctx *)op)->n;
        return 0;
    };
    hsize_t idx = 0;
    if (H5Literate(file, H5_INDEX_NAME, H5_ITER_NATIVE, &idx, cb, &ctx) < 0) return -1;
    return ctx.n;
}

// Particles of a step: the extent of its "x" dataset.
int64_t step_particles(hid_t grp) {
    Hid d(H5Dopen2(grp, "x", H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return -1;
    Hid s(H5Dget_space(d.id), H5Sclose);
    return (int64_t)H5Sget_simple_extent_npoints(s.id);
}

// The file attribute units of writeHeader (H5PartWrapperForPT.cpp:193-274), in its order.
const char *const kUnits[][2] = {
    {"idUnit", "1"}, {"xUnit", "m"}, {"yUnit", "m"}, {"zUnit", "m"},
    {"pxUnit", "#beta#gamma"}, {"pyUnit", "#beta#gamma"}, {"pzUnit", "#beta#gamma"},
    {"ptypeUnit", "1"}, {"qUnit", "C"}, {"TaitBryantAnglesUnit", "rad"},
    {"SPOSUnit", "m"}, {"TIMEUnit", "s"}, {"#gammaUnit", "1"}, {"ENERGYUnit", "MeV"},
    {"#varepsilonUnit", "m rad"}, {"#varepsilon-geomUnit", "m rad"}, {"#sigmaUnit", "1"},
    {"RMSXUnit", "m"}, {"centroidUnit", "m"}, {"RMSPUnit", "#beta#gamma"}, {"MEANPUnit", "#beta#gamma"},
    {"minXUnit", "m"}, {"maxXUnit", "m"}, {"minPUnit", "#beta#gamma"}, {"maxPUnit", "#beta#gamma"},
    {"dEUnit", "MeV"}, {"MASSUnit", "GeV"}, {"CHARGEUnit", "C"}, {"E-refUnit", "MV/m"}, {"B-refUnit", "T"},
    {"StepUnit", "1"}, {"LocalTrackStepUnit", "1"}, {"GlobalTrackStepUnit", "1"}, {"NumBunchUnit", "1"},
    {"NumPartUnit", "1"}, {"RefPartRUnit", "m"}, {"RefPartPUnit", "#beta#gamma"}, {"SteptoLastInjUnit", "1"},
    {"dump frequencyUnit", "1"}, {"dPhiGlobalUnit", "rad"}};

}  // namespace

extern "C" const char *opalsc_h5_last_error(void) { return g_err.c_str(); }

extern "C" int opalsc_h5_create(const char *path, const char *opal_version, int64_t dump_frequency,
                                double dphi_global) {
    if (!path) return fail(OPALSC_H5_EINVAL, "null path");
    Quiet q;
    Hid f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(OPALSC_H5_EIO, std::string("cannot create ") + path);
    bool ok = write_string_attr(f.id, "OPAL_version", opal_version ? opal_version : "OPAL (opalsc)");
    for (const auto &u : kUnits) ok = ok && write_string_attr(f.id, u[0], u[1]);
    ok = ok && write_i64(f.id, "dump frequency", dump_frequency);
    ok = ok && write_f64(f.id, "dPhiGlobal", &dphi_global, 1);
    return ok ? OPALSC_H5_OK : fail(OPALSC_H5_EIO, "writing the file attributes failed");
}

extern "C" int opalsc_h5_write_step(const char *path, int64_t N, const double *R, const double *P,
                                    const double *Q, const int64_t *id, const int32_t *ptype,
                                    const opalsc_h5_step *a, int64_t *step_out) {
    if (!path || !a || N < 0 || (N > 0 && (!R || !P || !Q || !id)))
        return fail(OPALSC_H5_EINVAL, "null argument or negative particle count");
    if (N == 0) return OPALSC_H5_OK;  // writeStep: nothing for an empty bunch
    Quiet q;
    Hid f(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(OPALSC_H5_EIO, std::string("cannot open ") + path);
    const int64_t k = count_steps(f.id);
    if (k < 0) return fail(OPALSC_H5_EIO, "cannot list the steps");
    Hid g(H5Gcreate2(f.id, step_name(k).c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(OPALSC_H5_EIO, "cannot create " + step_name(k));
    // writeStepHeader, H5PartWrapperForPT.cpp:326-373, in its order
    bool ok = write_string_attr(g.id, "OPAL_flavour", "opal-t");
    ok = ok && write_f64(g.id, "SPOS", &a->spos, 1) && write_f64(g.id, "RefPartR", a->ref_r, 3) &&
         write_f64(g.id, "centroid", a->centroid, 3) && write_f64(g.id, "RMSX", a->rms_x, 3) &&
         write_f64(g.id, "RefPartP", a->ref_p, 3) && write_f64(g.id, "MEANP", a->mean_p, 3) &&
         write_f64(g.id, "RMSP", a->rms_p, 3) && write_f64(g.id, "TaitBryantAngles", a->tait_bryant, 3) &&
         write_f64(g.id, "#varepsilon", a->eps_norm, 3) && write_f64(g.id, "#varepsilon-geom", a->eps_geom, 3) &&
         write_f64(g.id, "minX", a->min_x, 3) && write_f64(g.id, "maxX", a->max_x, 3) &&
         write_f64(g.id, "minP", a->min_p, 3) && write_f64(g.id, "maxP", a->max_p, 3) &&
         write_i64(g.id, "Step", k) && write_i64(g.id, "LocalTrackStep", a->local_track_step) &&
         write_i64(g.id, "GlobalTrackStep", a->global_track_step) && write_f64(g.id, "#sigma", &a->sigma, 1) &&
         write_f64(g.id, "TIME", &a->time, 1) && write_f64(g.id, "ENERGY", &a->energy, 1) &&
         write_f64(g.id, "dE", &a->dE, 1) && write_f64(g.id, "MASS", &a->mass, 1) &&
         write_f64(g.id, "CHARGE", &a->charge, 1) && write_i64(g.id, "NumBunch", a->num_bunch) &&
         write_i64(g.id, "SteptoLastInj", a->step_to_last_inj) && write_f64(g.id, "B-ref", a->B_ref, 3) &&
         write_f64(g.id, "E-ref", a->E_ref, 3);
    if (!ok) return fail(OPALSC_H5_EIO, "writing the step attributes failed");
    // writeStepData, H5PartWrapperForPT.cpp:379-424: one dataset per quantity
    std::vector<double> buf((size_t)N);
    const char *names[] = {"x", "y", "z", "px", "py", "pz"};
    for (int c = 0; c < 6 && ok; ++c) {
        const double *src = c < 3 ? R : P;
        for (int64_t i = 0; i < N; ++i) buf[(size_t)i] = src[3 * i + (c % 3)];
        ok = write_dataset(g.id, names[c], H5T_NATIVE_DOUBLE, buf.data(), (hsize_t)N);
    }
    ok = ok && write_dataset(g.id, "q", H5T_NATIVE_DOUBLE, Q, (hsize_t)N);
    ok = ok && write_dataset(g.id, "id", H5T_NATIVE_INT64, id, (hsize_t)N);
    std::vector<int32_t> pt;
    if (!ptype) pt.assign((size_t)N, 0);
    ok = ok && write_dataset(g.id, "ptype", H5T_NATIVE_INT32, ptype ? ptype : pt.data(), (hsize_t)N);
    if (!ok) return fail(OPALSC_H5_EIO, "writing the particle data failed");
    if (step_out) *step_out = k;
    return OPALSC_H5_OK;
}

extern "C" int opalsc_h5_num_steps(const char *path, int64_t *nsteps) {
    if (!path || !nsteps) return fail(OPALSC_H5_EINVAL, "null argument");
    Quiet q;
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(OPALSC_H5_EIO, std::string("cannot open ") + path);
    const int64_t k = count_steps(f.id);
    if (k < 0) return fail(OPALSC_H5_EIO, "cannot list the steps");
    *nsteps = k;
    return OPALSC_H5_OK;
}

static int open_step(const char *path, int64_t step, Hid &f, hid_t &grp, int64_t &k) {
    f.id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    f.close = H5Fclose;
    if (!f.ok()) return fail(OPALSC_H5_EIO, std::string("cannot open ") + path);
    const int64_t n = count_steps(f.id);
    if (n <= 0) return fail(OPALSC_H5_ERANGE, "the file holds no step");
    k = step < 0 ? n - 1 : step;
    if (k >= n) return fail(OPALSC_H5_ERANGE, "no step " + std::to_string((long long)k));
    grp = H5Gopen2(f.id, step_name(k).c_str(), H5P_DEFAULT);
    if (grp < 0) return fail(OPALSC_H5_EIO, "cannot open " + step_name(k));
    return OPALSC_H5_OK;
}

extern "C" int opalsc_h5_num_particles(const char *path, int64_t step, int64_t *N) {
    if (!path || !N) return fail(OPALSC_H5_EINVAL, "null argument");
    Quiet q;
    Hid f(-1, nullptr);
    hid_t grp = -1;
    int64_t k = 0;
    if (int e = open_step(path, step, f, grp, k)) return e;
    Hid g(grp, H5Gclose);
    const int64_t n = step_particles(g.id);
    if (n < 0) return fail(OPALSC_H5_EIO, "the step has no x dataset");
    *N = n;
    return OPALSC_H5_OK;
}

extern "C" int opalsc_h5_read_header(const char *path, int64_t *dump_frequency, double *dphi_global) {
    if (!path) return fail(OPALSC_H5_EINVAL, "null path");
    Quiet q;
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(OPALSC_H5_EIO, std::string("cannot open ") + path);
    // readHeader reads each only if the file holds it
    if (dump_frequency) (void)read_attr(f.id, "dump frequency", H5T_NATIVE_INT64, dump_frequency, 1);
    if (dphi_global) (void)read_attr(f.id, "dPhiGlobal", H5T_NATIVE_DOUBLE, dphi_global, 1);
    return OPALSC_H5_OK;
}

extern "C" int opalsc_h5_read_step(const char *path, int64_t step, int64_t first, int64_t last, double *R,
                                   double *P, double *Q, int64_t *id, opalsc_h5_step *a) {
    if (!path) return fail(OPALSC_H5_EINVAL, "null path");
    Quiet q;
    Hid f(-1, nullptr);
    hid_t grp = -1;
    int64_t k = 0;
    if (int e = open_step(path, step, f, grp, k)) return e;
    Hid g(grp, H5Gclose);
    if (a) {
        std::memset(a, 0, sizeof(*a));
        a->step = k;
        const struct {
            const char *name;
            double *v;
            int n;
        } dv[] = {{"SPOS", &a->spos, 1},         {"TIME", &a->time, 1},
                  {"RefPartR", a->ref_r, 3},     {"RefPartP", a->ref_p, 3},
                  {"centroid", a->centroid, 3},  {"RMSX", a->rms_x, 3},
                  {"MEANP", a->mean_p, 3},       {"RMSP", a->rms_p, 3},
                  {"TaitBryantAngles", a->tait_bryant, 3}, {"#varepsilon", a->eps_norm, 3},
                  {"#varepsilon-geom", a->eps_geom, 3},    {"minX", a->min_x, 3},
                  {"maxX", a->max_x, 3},         {"minP", a->min_p, 3},
                  {"maxP", a->max_p, 3},         {"#sigma", &a->sigma, 1},
                  {"ENERGY", &a->energy, 1},     {"dE", &a->dE, 1},
                  {"MASS", &a->mass, 1},         {"CHARGE", &a->charge, 1},
                  {"B-ref", a->B_ref, 3},        {"E-ref", a->E_ref, 3}};
        for (const auto &e : dv) (void)read_attr(g.id, e.name, H5T_NATIVE_DOUBLE, e.v, (hsize_t)e.n);
        const struct {
            const char *name;
            int64_t *v;
        } iv[] = {{"LocalTrackStep", &a->local_track_step}, {"GlobalTrackStep", &a->global_track_step},
                  {"NumBunch", &a->num_bunch}, {"SteptoLastInj", &a->step_to_last_inj}};
        for (const auto &e : iv) (void)read_attr(g.id, e.name, H5T_NATIVE_INT64, e.v, 1);
        // readStepHeader requires these
        const char *need[] = {"TIME", "SPOS", "LocalTrackStep", "GlobalTrackStep", "RefPartR", "RefPartP",
                              "TaitBryantAngles"};
        for (const char *n : need)
            if (H5Aexists(g.id, n) <= 0) return fail(OPALSC_H5_EIO, std::string("step attribute ") + n + " missing");
    }
    if (!R && !P && !Q && !id) return OPALSC_H5_OK;
    const int64_t n = step_particles(g.id);
    if (n < 0) return fail(OPALSC_H5_EIO, "the step has no x dataset");
    // readStepData, H5PartWrapperForPT.cpp:133-138
    if (last >= n || first > last || first < 0)
        return fail(OPALSC_H5_ERANGE, "the provided particle numbers don't match the number of particles in the file");
    const hsize_t cnt = (hsize_t)(last - first + 1), f0 = (hsize_t)first;
    std::vector<double> buf(cnt);
    const char *names[] = {"x", "y", "z", "px", "py", "pz"};
    for (int c = 0; c < 6; ++c) {
        double *dst = c < 3 ? R : P;
        if (!dst) continue;
        if (!read_slab(g.id, names[c], H5T_NATIVE_DOUBLE, f0, cnt, buf.data()))
            return fail(OPALSC_H5_EIO, std::string("cannot read ") + names[c]);
        for (hsize_t i = 0; i < cnt; ++i) dst[3 * i + (c % 3)] = buf[i];
    }
    if (Q && !read_slab(g.id, "q", H5T_NATIVE_DOUBLE, f0, cnt, Q)) return fail(OPALSC_H5_EIO, "cannot read q");
    if (id && !read_slab(g.id, "id", H5T_NATIVE_INT64, f0, cnt, id)) return fail(OPALSC_H5_EIO, "cannot read id");
    return OPALSC_H5_OK;
}
