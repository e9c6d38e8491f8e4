// mcm_output.h -- host side of the output path (libmcm_gpu.so): Fortran
// edit-descriptor formatting and the VTK particle writer.
//
// The writer produces the legacy ASCII VTK file of the reference's
// WriteVTK / WriteVTKmat (mcm_VTK.f90:1-200, 251-507) from the packed output
// fields (x, v, p, sigma per particle): same header, point, cell, cell-type
// and point-data blocks in the same order and with the same number formats
// (F15.8 for coordinates and velocities, E24.5 for scalars, the von Mises
// stress rounded to single precision as its REAL variable does, I5 for the
// particle type).  Differences, both deliberate:
//   - particle counts are written in full: the reference writes them with
//     I1..I5 (mcm_VTK.f90:56-85, 213-215) and its cell lines with I6, which
//     overflow to asterisks from 100,000 (counts) and 1,000,000 (cell ids)
//     particles and make the file unreadable;
//   - the per-material file's PTYPE block holds the cap model's mode of the
//     last update (mcm_f3dm25.f90 sets par%ptype) and 0 for particles of other
//     models, where the reference leaves ptype uninitialised
//     (mcm_database.f90:196).
// List-directed lines ("POINTS n double", "CELLS n 2n") are written with
// single blanks; VTK readers split on white space.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

namespace vtkout {

// Fortran Fw.d output of v (round to nearest; a value that does not fit in w
// columns is w asterisks; signed zeros and values that round to zero keep
// their sign; Inf / NaN right-justified).
inline std::string fmt_f(double v, int w, int d) {
    char buf[512];
    if (std::isnan(v)) snprintf(buf, sizeof buf, "NaN");
    else if (std::isinf(v)) snprintf(buf, sizeof buf, v > 0 ? "Inf" : "-Inf");
    else snprintf(buf, sizeof buf, "%.*f", d, v);
    std::string s(buf);
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// Fortran Ew.d output: [-]0.d1..dd E+xx (exponent field E+xx for |x| <= 99,
// +xxx without the E for |x| <= 999), right-justified in w columns.
inline std::string fmt_e(double v, int w, int d) {
    std::string s;
    if (std::isnan(v)) s = "NaN";
    else if (std::isinf(v)) s = v > 0 ? "Inf" : "-Inf";
    else {
        const bool neg = std::signbit(v);
        int ex = 0;
        std::string digits;
        if (v == 0.0) {
            digits.assign(d, '0');
        } else {
            char buf[64];
            snprintf(buf, sizeof buf, "%.*e", d - 1, std::fabs(v));   // d.dddde+xx: correctly rounded
            const char* e = strchr(buf, 'e');
            for (const char* p = buf; p < e; p++)
                if (*p >= '0' && *p <= '9') digits += *p;
            ex = atoi(e + 1) + 1;
        }
        char eb[16];
        if (ex >= -99 && ex <= 99) snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
        else snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
        s = std::string(neg ? "-" : "") + "0." + digits + eb;
    }
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// Fortran Iw output
inline std::string fmt_i(long v, int w) {
    std::string s = std::to_string(v);
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// Iw where the value fits, else the full number (the writer's count fields)
inline std::string fmt_i_full(long v, int w) {
    std::string s = std::to_string(v);
    if ((int)s.size() >= w) return s;
    return std::string(w - s.size(), ' ') + s;
}

// Packed output fields per particle (mcm_gpu_output_fields): x(3), v(3), p,
// sigma 11 22 33 12 13 23, e, efps, rho, h.
constexpr int kWords = 17;
enum { OX = 0, OV = 3, OP = 6, OS11 = 7, OS22 = 8, OS33 = 9, OS12 = 10, OS13 = 11, OS23 = 12 };

// Writes the VTK file of the particles whose 1-based material equals
// material (material 0: every particle, the layout of WriteVTK; material > 0:
// WriteVTKmat, with its PTYPE block).  mat may be null when material is 0;
// ptype (the cap model's mode per particle) may be null: all zero.
inline int write_vtk(const char* path, long np, const double* f, const int32_t* mat, int material,
                     const int32_t* ptype = nullptr) {
    FILE* out = fopen(path, "w");
    if (!out) return -1;
    auto take = [&](long i) { return material == 0 || (mat && mat[i] == material); };
    long n = 0;
    for (long i = 0; i < np; i++)
        if (take(i)) n++;
    std::string line;
    auto put = [&](const std::string& s) { fputs(s.c_str(), out); fputc('\n', out); };
    put("# vtk DataFile Version 3.0");
    put("TEST");
    put("ASCII");
    put("DATASET UNSTRUCTURED_GRID");
    put("POINTS " + std::to_string(n) + " double");
    for (long i = 0; i < np; i++)
        if (take(i)) {
            const double* r = f + (size_t)i * kWords;
            put(fmt_f(r[OX], 15, 8) + fmt_f(r[OX + 1], 15, 8) + fmt_f(r[OX + 2], 15, 8));
        }
    put("CELLS " + std::to_string(n) + " " + std::to_string(2 * n));
    for (long k = 0; k < n; k++) put("1 " + fmt_i_full(k, 6));
    put("CELL_TYPES " + fmt_i_full(n, 5));
    for (long k = 0; k < n; k++) put("1");
    put("POINT_DATA " + fmt_i_full(n, 5));
    put("VECTORS Velocities double");
    for (long i = 0; i < np; i++)
        if (take(i)) {
            const double* r = f + (size_t)i * kWords;
            put(fmt_f(r[OV], 15, 8) + fmt_f(r[OV + 1], 15, 8) + fmt_f(r[OV + 2], 15, 8));
        }
    static const char* names[7] = {"Pressure", "Stress_X", "Stress_Y", "Stress_Z", "Shear_Stress_XY",
                                   "Shear_Stress_XZ", "Shear_Stress_YZ"};
    static const int words[7] = {OP, OS11, OS22, OS33, OS12, OS13, OS23};
    for (int k = 0; k < 7; k++) {
        put(std::string("SCALARS ") + names[k] + " double 1");
        put("LOOKUP_TABLE default");
        for (long i = 0; i < np; i++)
            if (take(i)) put(fmt_e(f[(size_t)i * kWords + words[k]], 24, 5));
    }
    put("SCALARS VM_Stress double 1");
    put("LOOKUP_TABLE default");
    for (long i = 0; i < np; i++)
        if (take(i)) {
            const double* r = f + (size_t)i * kWords;
            const double a = r[OS11] - r[OS22], b = r[OS22] - r[OS33], c = r[OS33] - r[OS11];
            const double t = r[OS12] * r[OS12] + r[OS23] * r[OS23] + r[OS13] * r[OS13];
            const float vm = (float)std::sqrt(0.5 * (a * a + b * b + c * c + 6.0 * t));
            put(fmt_e((double)vm, 24, 5));
        }
    if (material > 0) {
        put("SCALARS PTYPE double 1");
        put("LOOKUP_TABLE default");
        for (long i = 0; i < np; i++)
            if (take(i)) put(fmt_i(ptype ? ptype[i] : 0, 5));
    }
    return fclose(out) == 0 ? 0 : -1;
}

}  // namespace vtkout
