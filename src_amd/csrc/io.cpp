// Host-side file formats around the space-charge step: FROMFILE particle
// distributions (read) and the DataSink statistics file (write).
#include "../../include/opalsc.h"
#include "kernels.hpp"

#include <cctype>
#include <cmath>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <limits>
#include <regex>
#include <sstream>
#include <string>

using opalsc::set_global_error;

namespace {

int io_fail(const std::string &msg) {
    set_global_error(msg);
    return OPALSC_EINVAL;
}

// Distribution::getNumberOfParticlesInFile (src/Distribution/Distribution.cpp:1073-1100):
// lines that are empty once '#' comments are removed are skipped; the first
// other line must start with a positive count.  The stream is left after it.
int read_count(std::ifstream &in, const std::string &path, int64_t &count) {
    static const std::regex comment("[[:space:]]*#.*");
    std::string line;
    do {
        std::getline(in, line);
        line = std::regex_replace(line, comment, "");
    } while (line.length() == 0 && !in.fail());
    std::stringstream ls(line);
    long n = 0;
    ls >> n;
    if (n <= 0)
        return io_fail("The file '" + path + "' does not seem to be an ASCII file containing a distribution.");
    count = (int64_t)n;
    return OPALSC_OK;
}

int open_input(std::ifstream &in, const char *path) {
    if (!path) return io_fail("null path");
    in.open(path);
    if (in.fail())
        return io_fail(std::string("Open file operation failed, please check if \"") + path + "\" really exists.");
    return OPALSC_OK;
}

// Distribution::converteVToBetaGamma (Distribution.cpp:1047-1054)
double ev_to_beta_gamma(double v, double mass_eV) {
    double value = std::copysign(std::sqrt(std::pow(std::abs(v) / mass_eV + 1.0, 2.0) - 1.0), v);
    if (std::abs(value) < std::numeric_limits<double>::epsilon())
        value = std::copysign(std::sqrt(2 * std::abs(v) / mass_eV), v);
    return value;
}

}  // namespace

extern "C" int opalsc_fromfile_count(const char *path, int64_t *N) {
    if (!N) return io_fail("null argument");
    std::ifstream in;
    if (int e = open_input(in, path)) return e;
    return read_count(in, path, *N);
}

// Distribution::createDistributionFromFile (Distribution.cpp:1103-1250) on one
// rank: x px y py z pz per particle read with operator>>; as there, a value the
// stream cannot read leaves the particle's remaining components 0.
extern "C" int opalsc_fromfile_read(const char *path, int64_t N, int momenta_in_ev, double mass_eV, double *R,
                                    double *P) {
    if ((N > 0 && (!R || !P)) || N < 0) return io_fail("null argument");
    if (momenta_in_ev && !(mass_eV > 0.0)) return io_fail("mass_eV must be positive for momenta in eV");
    std::ifstream in;
    if (int e = open_input(in, path)) return e;
    int64_t count = 0;
    if (int e = read_count(in, path, count)) return e;
    if (count != N)
        return io_fail("the file holds " + std::to_string(count) + " particles, not " + std::to_string(N));
    for (int64_t i = 0; i < N; ++i) {
        double r[3] = {0.0, 0.0, 0.0}, p[3] = {0.0, 0.0, 0.0};
        in >> r[0] >> p[0] >> r[1] >> p[1] >> r[2] >> p[2];
        if (momenta_in_ev)
            for (int d = 0; d < 3; ++d) p[d] = ev_to_beta_gamma(p[d], mass_eV);
        for (int d = 0; d < 3; ++d) {
            R[3 * i + d] = r[d];
            P[3 * i + d] = p[d];
        }
    }
    return OPALSC_OK;
}

namespace {

struct Column {
    const char *name, *type, *units, *desc;
};

// DataSink::writeSDDSHeader (src/Structure/DataSink.cpp:585-1010): the columns
// in file order (number, name, type, units, description as written there).
const Column kColumns[] = {
    {"t", "double", "ns", "1 Time"},
    {"s", "double", "m", "2 Average Longitudinal Position"},
    {"numParticles", "long", "1", "3 Number of Macro Particles"},
    {"charge", "double", "1", "4 Bunch Charge"},
    {"energy", "double", "MeV", "5 Mean Energy"},
    {"rms_x", "double", "m", "6 RMS Beamsize in x"},
    {"rms_y", "double", "m", "7 RMS Beamsize in y"},
    {"rms_s", "double", "m", "8 RMS Beamsize in s"},
    {"rms_px", "double", "1", "9 RMS Momenta in x"},
    {"rms_py", "double", "1", "10 RMS Momenta in y"},
    {"rms_ps", "double", "1", "11 RMS Momenta in s"},
    {"emit_x", "double", "m", "12 Normalized Emittance x"},
    {"emit_y", "double", "m", "13 Normalized Emittance y"},
    {"emit_s", "double", "m", "14 Normalized Emittance s"},
    {"mean_x", "double", "m", "15 Mean Beam Position in x"},
    {"mean_y", "double", "m", "16 Mean Beam Position in y"},
    {"mean_s", "double", "m", "17 Mean Beam Position in s"},
    {"ref_x", "double", "m", "18 x coordinate of reference particle in lab cs"},
    {"ref_y", "double", "m", "19 y coordinate of reference particle in lab cs"},
    {"ref_z", "double", "m", "20 z coordinate of reference particle in lab cs"},
    {"ref_px", "double", "1", "21 x momentum of reference particle in lab cs"},
    {"ref_py", "double", "1", "22 y momentum of reference particle in lab cs"},
    {"ref_pz", "double", "1", "23 z momentum of reference particle in lab cs"},
    {"max_x", "double", "m", "24 Max Beamsize in x"},
    {"max_y", "double", "m", "25 Max Beamsize in y"},
    {"max_s", "double", "m", "26 Max Beamsize in s"},
    {"xpx", "double", "1", "27 Correlation xpx"},
    {"ypy", "double", "1", "28 Correlation ypyy"},
    {"zpz", "double", "1", "29 Correlation zpz"},
    {"Dx", "double", "m", "30 Dispersion in x"},
    {"DDx", "double", "1", "31 Derivative of dispersion in x"},
    {"Dy", "double", "m", "32 Dispersion in y"},
    {"DDy", "double", "1", "33 Derivative of dispersion in y"},
    {"Bx_ref", "double", "T", "34 Bx-Field component of ref particle"},
    {"By_ref", "double", "T", "35 By-Field component of ref particle"},
    {"Bz_ref", "double", "T", "36 Bz-Field component of ref particle"},
    {"Ex_ref", "double", "MV/m", "37 Ex-Field component of ref particle"},
    {"Ey_ref", "double", "MV/m", "38 Ey-Field component of ref particle"},
    {"Ez_ref", "double", "MV/m", "39 Ez-Field component of ref particle"},
    {"dE", "double", "MeV", "40 energy spread of the beam"},
    {"dt", "double", "ns", "41 time step size"},
    {"partsOutside", "double", "1", "42 outside n*sigma of the beam"},
};
const Column kParticle0Columns[] = {
    {"R0_x", "double", "m", "43 R0 Particle position in x"},
    {"R0_y", "double", "m", "44 R0 Particle position in y"},
    {"R0_s", "double", "m", "45 R0 Particle position in z"},
    {"P0_x", "double", "1", "46 R0 Particle momentum in x"},
    {"P0_y", "double", "1", "47 R0 Particle momentum in y"},
    {"P0_s", "double", "1", "48 R0 Particle momentum in z"},
};

void write_column(std::ostream &os, const Column &c) {
    const char *indent = "        ";
    os << "&column\n"
       << indent << "name=" << c.name << ",\n"
       << indent << "type=" << c.type << ",\n"
       << indent << "units=" << c.units << ",\n"
       << indent << "description=\"" << c.desc << "\"\n"
       << "&end\n";
}

void write_header(std::ostream &os, const char *input_fn, bool particle0) {
    const char *indent = "        ";
    char date[12] = {0}, tim[12] = {0};
    const std::time_t now = std::time(nullptr);
    std::strftime(date, sizeof(date), "%d/%m/%Y", std::localtime(&now));
    std::strftime(tim, sizeof(tim), "%H:%M:%S", std::localtime(&now));
    os << "SDDS1" << std::endl;
    os << "&description\n"
       << indent << "text=\"Statistics data '" << (input_fn ? input_fn : "") << "' " << date << " " << tim
       << "\",\n"
       << indent << "contents=\"stat parameters\"\n"
       << "&end\n";
    os << "&parameter\n"
       << indent << "name=processors,\n"
       << indent << "type=long,\n"
       << indent << "description=\"Number of Cores used\"\n"
       << "&end\n";
    os << "&parameter\n"
       << indent << "name=revision,\n"
       << indent << "type=string,\n"
       << indent << "description=\"git revision of opal\"\n"
       << "&end\n";
    os << "&parameter\n"
       << indent << "name=flavor,\n"
       << indent << "type=string,\n"
       << indent << "description=\"OPAL flavor that wrote file\"\n"
       << "&end\n";
    for (const Column &c : kColumns) write_column(os, c);
    if (particle0)
        for (const Column &c : kParticle0Columns) write_column(os, c);
    os << "&data\n"
       << indent << "mode=ascii,\n"
       << indent << "no_row_counts=1\n"
       << "&end\n";
    os << 1 << std::endl;
    os << "opalsc (MI355X space-charge step) ABI " << OPALSC_ABI_VERSION << std::endl;
    os << "opal-t" << std::endl;
}

}  // namespace

// DataSink::doWriteStatData (DataSink.cpp:281-430): each value followed by a
// tab padded to width 10, doubles in scientific notation with 15 digits.
extern "C" int opalsc_stat_write(const char *path, const char *input_fn, int new_file, const opalsc_stat_row *row) {
    if (!path || !row) return io_fail("null argument");
    std::ofstream os(path, new_file ? std::ios::out : std::ios::app);
    if (!os) return io_fail(std::string("cannot open '") + path + "' for writing");
    os.precision(15);
    os.setf(std::ios::scientific, std::ios::floatfield);
    if (new_file) write_header(os, input_fn, row->has_particle0 != 0);
    const int pwi = 10;
    auto put = [&](auto v) { os << v << std::setw(pwi) << "\t"; };
    put(row->t_ns);
    put(row->s);
    put(row->num_particles);
    put(row->charge);
    put(row->energy_MeV);
    for (int d = 0; d < 3; ++d) put(row->rms_r[d]);
    for (int d = 0; d < 3; ++d) put(row->rms_p[d]);
    for (int d = 0; d < 3; ++d) put(row->emit[d]);
    for (int d = 0; d < 3; ++d) put(row->mean_r[d]);
    for (int d = 0; d < 3; ++d) put(row->ref_r[d]);
    for (int d = 0; d < 3; ++d) put(row->ref_p[d]);
    for (int d = 0; d < 3; ++d) put(row->max_r[d]);
    for (int d = 0; d < 3; ++d) put(row->corr[d]);
    put(row->Dx);
    put(row->DDx);
    put(row->Dy);
    put(row->DDy);
    for (int d = 0; d < 3; ++d) put(row->B_ref[d]);
    for (int d = 0; d < 3; ++d) put(row->E_ref[d]);
    put(row->dE_MeV);
    put(row->dt_ns);
    put(row->parts_outside);
    if (row->has_particle0) {
        for (int d = 0; d < 3; ++d) put(row->R0[d]);
        for (int d = 0; d < 3; ++d) put(row->P0[d]);
    }
    os << std::endl;
    if (!os) return io_fail(std::string("write to '") + path + "' failed");
    return OPALSC_OK;
}

// ---- 2D field maps ------------------------------------------------------------
namespace opalsc {

namespace {

// Fieldmap::getLine (src/Classic/Fields/Fieldmap.cpp:519-540): the next line
// with an alphanumeric character once a '#' comment is cut, trimmed to its
// first and last such character.
const char *kAlnum = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789.-+\211";

bool get_line(std::ifstream &in, std::string &buffer) {
    size_t firstof = 0, lastof;
    do {
        if (!std::getline(in, buffer)) {
            buffer.clear();
            return false;
        }
        buffer = buffer.substr(0, buffer.find('#'));
        lastof = buffer.find_last_of(kAlnum);
        firstof = buffer.find_first_of(kAlnum);
    } while (lastof == std::string::npos);
    buffer = buffer.substr(firstof, lastof - firstof + 1);
    return true;
}

// Fieldmap::interpreteLine: the line must hold exactly the values asked for.
template <typename... T>
bool interprete_line(std::ifstream &in, T &...values) {
    std::string buffer;
    if (!get_line(in, buffer)) return false;
    std::istringstream interpreter(buffer);
    bool ok = true;
    ((ok = ok && static_cast<bool>(interpreter >> values)), ...);
    if (!ok) return false;
    std::string rest;
    return !(interpreter >> rest);
}

// Fieldmap::interpreteEOF (Fieldmap.cpp:542-556): nothing but comments left.
bool only_comments_left(std::ifstream &in) {
    std::string line;
    while (std::getline(in, line)) {
        line = line.substr(0, line.find('#'));
        if (line.find_first_of(kAlnum) != std::string::npos) return false;
    }
    return true;
}

}  // namespace

// The constructors and readMap of FM2DMagnetoStatic (FM2DMagnetoStatic.cpp:11-140)
// and FM2DDynamic (FM2DDynamic.cpp:13-150): first line "<type> XZ|ZX" (the type
// from its first four characters, Fieldmap.cpp:318-330); then z and r ranges in
// cm with the number of spacings (r first for ZX), a dynamic map's frequency
// (MHz) between them; then (nz+1)(nr+1) lines of 2 (magnetostatic: Bz Br, or
// Br Bz for ZX) or 4 values (dynamic: Ez Er - Bt, or Er Ez Bt - for ZX),
// z fastest for XZ, r fastest for ZX.  Fields are normalised to the largest
// on-axis |Bz| / |Ez| (the default of Fieldmap::normalize_m); a dynamic map's E
// is converted from MV/m, its H to B with mu_0.  A third word on the first
// line makes the reference's parse fail (it overwrites the orientation), and
// so it does here.
int parse_fieldmap_2d(const char *path, FieldMapHost &out, std::string &err) {
    std::ifstream file(path ? path : "");
    if (!path || !file.good()) {
        err = std::string("field map '") + (path ? path : "") + "' cannot be opened";
        return -1;
    }
    std::string line;
    if (!get_line(file, line)) {
        err = std::string("field map '") + path + "' is empty";
        return -1;
    }
    const std::string magic = line.substr(0, 4);
    int kind;
    if (magic == "2DMa") kind = 0;
    else if (magic == "2DDy") kind = 1;
    else {
        err = std::string("field map '") + path + "' is not a 2D magnetostatic or 2D dynamic map";
        return -1;
    }
    std::istringstream first(line);
    std::string w1, w2, w3;
    first >> w1 >> w2;
    if (first >> w3) {
        for (char &ch : w3) ch = (char)std::toupper((unsigned char)ch);
        if (w3 != "TRUE" && w3 != "FALSE")
            err = "The third string on the first line of 2D field maps has to be either TRUE or FALSE";
        else
            err = std::string("An error occured when reading the fieldmap '") + path + "'";
        return -1;
    }
    double zbegin = 0, zend = 0, rbegin = 0, rend = 0, freq = 0;
    int nz = 0, nr = 0;
    bool swap, ok = true;
    if (w2 == "ZX") {
        swap = true;
        ok = interprete_line(file, rbegin, rend, nr);
        if (kind == 1) ok = ok && interprete_line(file, freq);
        ok = ok && interprete_line(file, zbegin, zend, nz);
    } else if (w2 == "XZ") {
        swap = false;
        ok = interprete_line(file, zbegin, zend, nz);
        if (kind == 1) ok = ok && interprete_line(file, freq);
        ok = ok && interprete_line(file, rbegin, rend, nr);
    } else {
        err = std::string("An error occured when reading the fieldmap '") + path + "'";
        return -1;
    }
    if (!ok || nz < 1 || nr < 1) {
        err = std::string("An error occured when reading the fieldmap '") + path + "'";
        return -1;
    }
    const int gz = nz + 1, gr = nr + 1;
    out.a.assign((size_t)gz * gr, 0.0);
    out.b.assign((size_t)gz * gr, 0.0);
    out.c.assign(kind == 1 ? (size_t)gz * gr : 0, 0.0);
    for (int outer = 0; outer < (swap ? gz : gr) && ok; ++outer)
        for (int inner = 0; inner < (swap ? gr : gz) && ok; ++inner) {
            const int i = swap ? outer : inner, j = swap ? inner : outer;
            const size_t k = (size_t)i + (size_t)j * gz;
            double v1 = 0, v2 = 0, v3 = 0, v4 = 0;
            if (kind == 0) {
                ok = interprete_line(file, v1, v2);
                out.a[k] = swap ? v2 : v1;  // Bz
                out.b[k] = swap ? v1 : v2;  // Br
            } else {
                ok = interprete_line(file, v1, v2, v3, v4);
                out.a[k] = swap ? v2 : v1;  // Ez
                out.b[k] = swap ? v1 : v2;  // Er
                out.c[k] = swap ? v3 : v4;  // Bt (H)
            }
        }
    if (!ok || !only_comments_left(file)) {
        err = std::string("An error occured when reading the fieldmap '") + path + "'";
        return -1;
    }
    double maxv = 0.0;
    for (int i = 0; i < gz; ++i)
        if (std::abs(out.a[i]) > maxv) maxv = std::abs(out.a[i]);
    if (kind == 0) {
        for (size_t k = 0; k < out.a.size(); ++k) {
            out.a[k] /= maxv;
            out.b[k] /= maxv;
        }
    } else {
        const double mu_0 = 1.256637061e-06;  // src/Classic/Physics/Physics.h:55
        for (size_t k = 0; k < out.a.size(); ++k) {
            out.a[k] *= 1.e6 / maxv;
            out.b[k] *= 1.e6 / maxv;
            out.c[k] *= mu_0 / maxv;
        }
    }
    FieldMapDesc &d = out.d;
    d.kind = kind;
    rbegin /= 100.;
    rend /= 100.;
    zbegin /= 100.;
    zend /= 100.;
    d.hr = (rend - rbegin) / nr;
    d.hz = (zend - zbegin) / nz;
    d.nz = gz;
    d.nr = gr;
    d.zstart = zbegin;
    d.length = zend - zbegin;
    d.scale = 1.0;
    d.omega = 0.0;
    d.phase = 0.0;
    d.a = 0;
    d.b = (int64_t)out.a.size();
    d.c = kind == 1 ? 2 * (int64_t)out.a.size() : 0;
    out.frequency = freq * (2.0 * 3.14159265358979323846) * 1e6;
    return 0;
}

}  // namespace opalsc
