// The output writers of the DTFE command line, host code beside the drop-in:
//   writeOutputData ............ src/input_output.cc:213-298 (file names, field order)
//   format 101 raw binary ...... src/io/binary_io.cc:184-216
//   formats 111-114 text ....... src/io/text_io.cc:224-628
//   format 100 density header .. src/io/density_file_io.cc:29-304
// Text is written with the stream's default formatting (6 significant
// digits), as the reference's `outputFile << value` is.
#include "input_output.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

namespace dtfe_dropin {
[[noreturn]] void fatal(const std::string &msg);
}

namespace {

using dtfe_dropin::fatal;

// one row: a Real and its newline, or each component of a Pvector followed
// by a tab, then the newline (text_io.cc:235-236, :255-260)
inline void put_row(std::ostream &o, const Real &v) { o << v << "\n"; }
template <size_t N>
void put_row(std::ostream &o, const Pvector<Real, N> &v) {
    for (size_t j = 0; j < N; ++j) o << v[j] << "\t";
    o << "\n";
}

std::ofstream open_text(const std::string &name) {
    std::ofstream f(name.c_str(), std::ios::out);
    if (!f) fatal("The program could not open the output file '" + name + "' for writing.");
    return f;
}

std::ofstream open_binary(const std::string &name) {
    std::ofstream f(name.c_str(), std::ios::out | std::ios::binary);
    if (!f) fatal("The program could not open the output file '" + name + "' for writing.");
    return f;
}

void check(std::ostream &f, const std::string &name) {
    if (!f) fatal("The program could not write to the output file '" + name + "'.");
}

// writes data in blocks of at most 256^3 rows (binary_io.cc:196-207)
template <typename T>
void write_rows(std::ostream &f, const std::vector<T> &d) {
    const size_t block = 256 * 256 * 256;
    for (size_t at = 0; at < d.size(); at += block) {
        const size_t n = std::min(block, d.size() - at);
        f.write(reinterpret_cast<const char *>(d.data() + at), n * sizeof(T));
    }
}

template <typename T>
void write_101(const std::vector<T> &d, const std::string &name) {
    std::ofstream f = open_binary(name);
    write_rows(f, d);
    check(f, name);
}

template <typename T>
void write_111(const std::vector<T> &d, const std::string &name) {
    std::ofstream f = open_text(name);
    for (const T &v : d) put_row(f, v);
    check(f, name);
}

template <typename T>
void write_112(const std::vector<T> &d, const std::string &name, const User_options &o) {
    if (o.userDefinedSampling)
        fatal("You cannot use the function 'writeTextFile_gridIndex' to write the data to a text file when using "
              "user defined coordinates since there are no grid indices associated to each sampling point.");
    std::ofstream f = open_text(name);
    const size_t *g = &o.gridSize[0];
    for (size_t i = 0; i < g[0]; ++i)
        for (size_t j = 0; j < g[1]; ++j)
            for (size_t k = 0; k < g[2]; ++k) {
                f << i << "\t" << j << "\t" << k << "\t";
                put_row(f, d[i * g[1] * g[2] + j * g[2] + k]);
            }
    check(f, name);
}

template <typename T>
void write_113(const std::vector<T> &d, const std::string &name, const User_options &o) {
    if (o.userDefinedSampling || o.redshiftConeOn)
        fatal("You cannot use the function 'writeTextFile_samplingPosition' to write the data to a text file when "
              "using redshift cone or user defined coordinates since teh sampling coordinates are incorrect in "
              "this case.");
    std::ofstream f = open_text(name);
    const size_t *g = &o.gridSize[0];
    Real dx[3];
    for (int a = 0; a < 3; ++a) dx[a] = (o.region[2 * a + 1] - o.region[2 * a]) / g[a];
    for (size_t i = 0; i < g[0]; ++i)
        for (size_t j = 0; j < g[1]; ++j)
            for (size_t k = 0; k < g[2]; ++k) {
                Real x = o.region[0] + dx[0] * (i + 0.5);
                Real y = o.region[2] + dx[1] * (j + 0.5);
                Real z = o.region[4] + dx[2] * (k + 0.5);
                f << x << "\t" << y << "\t" << z << "\t";
                put_row(f, d[i * g[1] * g[2] + j * g[2] + k]);
            }
    check(f, name);
}

template <typename T>
void write_114(const std::vector<T> &d, const std::string &name, const User_options &o) {
    if (o.userDefinedSampling)
        fatal("You cannot use the function 'writeTextFile_redshiftConePosition' to write the data to a text file "
              "when using user defined coordinates since the program doesn't know how to compute the redshift "
              "cone coordinates.");
    if (!o.redshiftConeOn)
        fatal("The function 'writeTextFile_redshiftConePosition' can be used to write the data to a text file only "
              "when using redshift cone coordinates since it writes the redshift cone coordinates in the file too.");
    std::ofstream f = open_text(name);
    const size_t *g = &o.gridSize[0];
    Real dx[3];
    for (int a = 0; a < 3; ++a) dx[a] = (o.redshiftCone[2 * a + 1] - o.redshiftCone[2 * a]) / g[a];
    const Real factor = 3.14 / 180.;  // the writer's own degree conversion (text_io.cc:511)
    const double org[3] = {(double)o.originPosition[0], (double)o.originPosition[1], (double)o.originPosition[2]};
    for (size_t i = 0; i < g[0]; ++i)
        for (size_t j = 0; j < g[1]; ++j)
            for (size_t k = 0; k < g[2]; ++k) {
                Real r = o.redshiftCone[0] + dx[0] * (i + 0.5);
                Real theta = o.redshiftCone[2] + dx[1] * (j + 0.5);
                Real phi = o.redshiftCone[4] + dx[2] * (k + 0.5);
                // the C library's double sin / cos, as the unqualified calls there resolve
                Real x = org[0] + r * ::sin((double)(theta * factor)) * ::cos((double)(phi * factor));
                Real y = org[1] + r * ::sin((double)(theta * factor)) * ::sin((double)(phi * factor));
                Real z = org[2] + r * ::cos((double)(theta * factor));
                f << x << "\t" << y << "\t" << z << "\t";
                put_row(f, d[i * g[1] * g[2] + j * g[2] + k]);
            }
    check(f, name);
}

// Density_header of density_file_io.cc:44-160: 1024 bytes
constexpr int kFill = 1024 - 13 * 8 - 8 * 18 - 8 * 2;
struct DensityHeader {
    uint64_t gridSize[3];
    uint64_t totalGrid;
    int32_t fileType;
    int32_t noDensityFiles;
    int32_t densityFileGrid[3];
    int32_t indexDensityFile;
    double box[6];
    uint64_t npartTotal[6];
    double mass[6];
    double time, redshift, BoxSize, Omega0, OmegaLambda, HubbleParam;
    uint64_t method;
    char fill[kFill];
    uint64_t FILE_ID;
};
static_assert(sizeof(DensityHeader) == 1024, "the density header is 1024 bytes");

DensityHeader density_header(const User_options &o, const std::string &variable) {
    DensityHeader h;
    std::memset(&h, 0, sizeof(h));  // the reference leaves mass[] unset; zero here
    for (int i = 0; i < 3; ++i) h.densityFileGrid[i] = 1;
    h.fileType = -1;
    h.noDensityFiles = 1;
    h.indexDensityFile = -1;
    h.time = h.redshift = h.BoxSize = h.Omega0 = h.OmegaLambda = h.HubbleParam = -1.;
    h.method = (uint64_t)-1;
    h.FILE_ID = 1;
    // updateDensityHeader (:154-206)
    std::vector<Real> region(o.region.coords);
    if (o.regionOn && !o.regionMpcOn)
        for (size_t i = 0; i < region.size(); ++i)
            region[i] = o.boxCoordinates[i % 2] + region[i] * (o.boxCoordinates[i % 2 + 1] - o.boxCoordinates[i % 2]);
    for (int i = 0; i < 3; ++i) h.gridSize[i] = o.gridSize[i];
    if (o.partNo >= 0)
        for (int i = 0; i < 3; ++i) h.densityFileGrid[i] = (int32_t)o.partition[i];
    for (int i = 0; i < 6; ++i) h.box[i] = o.regionOn ? region[i] : o.boxCoordinates[i];
    h.totalGrid = h.gridSize[0] * h.gridSize[1] * h.gridSize[2];
    const size_t len = std::min<size_t>(o.programOptions.size(), kFill - 4);
    std::memcpy(h.fill, o.programOptions.data(), len);
    h.fill[len] = ' ', h.fill[len + 1] = ';', h.fill[len + 2] = ' ', h.fill[len + 3] = ' ';
    if (o.partNo >= 0) h.indexDensityFile = o.partNo;
    if (o.DTFE) h.method = 1;
    else if (o.TSC) h.method = 2;
    else if (o.SPH) h.method = 3;
    auto has = [&](const char *w) { return variable.find(w) != std::string::npos; };
    if (has("density")) h.fileType = 1;
    else if (has("velocity gradient")) h.fileType = 12;
    else if (has("velocity divergence")) h.fileType = 13;
    else if (has("velocity shear")) h.fileType = 14;
    else if (has("velocity vorticity")) h.fileType = 15;
    else if (has("velocity standard deviation")) h.fileType = 16;
    else if (has("velocity")) h.fileType = 11;
    else if (has("scalar")) h.fileType = 20;
    return h;
}

// writeSpecialFile (:256-298): [8-byte 1024][header][8-byte 1024][8-byte n][data][8-byte n]
template <typename T>
void write_100(const std::vector<T> &d, const std::string &name, const std::string &variable,
               const User_options &o) {
    DensityHeader h = density_header(o, variable);
    std::ofstream f = open_binary(name);
    uint64_t b = sizeof(h);
    f.write(reinterpret_cast<const char *>(&b), sizeof(b));
    f.write(reinterpret_cast<const char *>(&h), sizeof(h));
    f.write(reinterpret_cast<const char *>(&b), sizeof(b));
    b = d.size() * sizeof(T);
    f.write(reinterpret_cast<const char *>(&b), sizeof(b));
    write_rows(f, d);
    f.write(reinterpret_cast<const char *>(&b), sizeof(b));
    check(f, name);
}

// OutputData::write (input_output.cc:188-210)
template <typename T>
void write(const std::vector<T> &d, const std::string &name, const std::string &variable, const User_options &o) {
    switch (o.outputFileType) {
        case 101: write_101(d, name); break;
        case 111: write_111(d, name); break;
        case 112: write_112(d, name, o); break;
        case 113: write_113(d, name, o); break;
        case 114: write_114(d, name, o); break;
        case 100: write_100(d, name, variable, o); break;
        case 121: fatal("output file type 121 (the author's own format) is not provided by the MI355X drop-in");
        default: break;  // the reference writes nothing for other values
    }
}

}  // namespace

void writeOutputData(Quantities &u, Quantities &a, User_options const &o) {
    const std::string &f = o.outputFilename;
    if (o.uField.density) write(u.density, f + ".den", "density", o);
    if (o.aField.density) write(a.density, f + ".a_den", "volume averaged density", o);
    if (o.uField.velocity) write(u.velocity, f + ".vel", "velocity", o);
    if (o.aField.velocity) write(a.velocity, f + ".a_vel", "volume averaged velocity", o);
    if (o.uField.velocity_gradient) write(u.velocity_gradient, f + ".velGrad", "velocity gradient", o);
    if (o.aField.velocity_gradient)
        write(a.velocity_gradient, f + ".a_velGrad", "volume averaged velocity gradient", o);
    if (o.uField.velocity_divergence) write(u.velocity_divergence, f + ".velDiv", "velocity divergence", o);
    if (o.aField.velocity_divergence)
        write(a.velocity_divergence, f + ".a_velDiv", "volume averaged velocity divergence", o);
    if (o.uField.velocity_shear) write(u.velocity_shear, f + ".velShear", "velocity shear", o);
    if (o.aField.velocity_shear) write(a.velocity_shear, f + ".a_velShear", "volume averaged velocity shear", o);
    if (o.uField.velocity_vorticity) write(u.velocity_vorticity, f + ".velVort", "velocity vorticity", o);
    if (o.aField.velocity_vorticity)
        write(a.velocity_vorticity, f + ".a_velVort", "volume averaged velocity vorticity", o);
    if (o.aField.velocity_std)
        write(a.velocity_std, f + ".a_velStd", "volume averaged velocity standard deviation", o);
    if (o.uField.scalar) write(u.scalar, f + ".scalar", "scalar", o);
    if (o.aField.scalar) write(a.scalar, f + ".a_scalar", "volume averaged scalar", o);
    if (o.uField.scalar_gradient) write(u.scalar_gradient, f + ".scalarGrad", "scalar gradient", o);
    if (o.aField.scalar_gradient) write(a.scalar_gradient, f + ".a_scalarGrad", "volume averaged scalar gradient", o);
}
