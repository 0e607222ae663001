/*
 * LAMMPS dump I/O of libedmd_hip.so.
 *
 * Reader: the reference's load path (EDMD.c:987-995, 1387-1451) reads the
 * last frame through parser.c: the atom count from NUMBER OF ATOMS, Lx and
 * Ly from column 1 of the first two BOX rows (parser.c:444-479), and each
 * atom property from the first ATOMS column whose name starts with the
 * property name (parser.c:900-917), parsed as %lf.  This reader follows the
 * same rules.  Writer: the saveTXT layout (EDMD.c:5012-5153).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/edmd.h"

int edmd_internal_fail(int code, const char* msg);

namespace {

struct Frame {
    int n = -1;
    double Lx = 0, Ly = 0;
    std::vector<std::string> cols;
    long atoms_pos = -1; /* file offset of the first atom row */
};

bool read_file(const char* path, std::string& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, k);
    fclose(f);
    return true;
}

/* next line starting at pos; advances pos past the newline */
bool next_line(const std::string& s, size_t& pos, std::string& line) {
    if (pos >= s.size()) return false;
    size_t e = s.find('\n', pos);
    if (e == std::string::npos) e = s.size();
    line.assign(s, pos, e - pos);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    pos = e + 1;
    return true;
}

std::vector<std::string> tokens(const std::string& line) {
    std::vector<std::string> t;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) i++;
        size_t j = i;
        while (j < line.size() && line[j] != ' ' && line[j] != '\t') j++;
        if (j > i) t.push_back(line.substr(i, j - i));
        i = j;
    }
    return t;
}

bool parse_last_frame(const std::string& s, Frame& fr, size_t& atoms_at) {
    size_t last = s.rfind("ITEM: TIMESTEP");
    if (last == std::string::npos) return false;
    size_t pos = last;
    std::string line;
    while (next_line(s, pos, line)) {
        if (line.compare(0, 6, "ITEM: ") != 0) continue;
        if (line.compare(6, 6, "NUMBER") == 0) {
            if (!next_line(s, pos, line)) return false;
            fr.n = atoi(line.c_str());
        } else if (line.compare(6, 3, "BOX") == 0) {
            if (!next_line(s, pos, line)) return false;
            std::vector<std::string> t = tokens(line);
            if (t.size() < 2) return false;
            fr.Lx = strtod(t[1].c_str(), nullptr);
            if (!next_line(s, pos, line)) return false;
            t = tokens(line);
            if (t.size() < 2) return false;
            fr.Ly = strtod(t[1].c_str(), nullptr);
        } else if (line.compare(6, 5, "ATOMS") == 0) {
            std::vector<std::string> t = tokens(line);
            fr.cols.assign(t.begin() + 2, t.end());
            atoms_at = pos;
            return fr.n > 0;
        }
    }
    return false;
}

int column(const Frame& fr, const char* prop) {
    size_t k = strlen(prop);
    for (size_t c = 0; c < fr.cols.size(); c++)
        if (fr.cols[c].compare(0, k, prop) == 0) return (int)c;
    return -1;
}

}  // namespace

extern "C" int edmd_dump_info(const char* path, int* n, double* Lx, double* Ly) {
    std::string s;
    if (!path || !read_file(path, s)) return edmd_internal_fail(EDMD_ERR_IO, "cannot read dump file");
    Frame fr;
    size_t at = 0;
    if (!parse_last_frame(s, fr, at)) return edmd_internal_fail(EDMD_ERR_IO, "malformed dump file");
    if (n) *n = fr.n;
    if (Lx) *Lx = fr.Lx;
    if (Ly) *Ly = fr.Ly;
    return EDMD_OK;
}

extern "C" int edmd_dump_read(const char* path, int n, double* x, double* y, double* vx, double* vy,
                              double* rad, double* m) {
    std::string s;
    if (!path || !read_file(path, s)) return edmd_internal_fail(EDMD_ERR_IO, "cannot read dump file");
    Frame fr;
    size_t pos = 0;
    if (!parse_last_frame(s, fr, pos)) return edmd_internal_fail(EDMD_ERR_IO, "malformed dump file");
    if (n > fr.n) return edmd_internal_fail(EDMD_ERR_ARG, "dump holds fewer atoms than requested");
    const char* names[6] = {"x", "y", "vx", "vy", "radius", "m"};
    double* dst[6] = {x, y, vx, vy, rad, m};
    int col[6];
    for (int k = 0; k < 6; k++) {
        col[k] = column(fr, names[k]);
        if (dst[k] && col[k] < 0) return edmd_internal_fail(EDMD_ERR_IO, "dump lacks an atom column");
    }
    std::string line;
    for (int i = 0; i < n; i++) {
        if (!next_line(s, pos, line)) return edmd_internal_fail(EDMD_ERR_IO, "dump ends early");
        std::vector<std::string> t = tokens(line);
        for (int k = 0; k < 6; k++) {
            if (!dst[k]) continue;
            if ((size_t)col[k] >= t.size()) return edmd_internal_fail(EDMD_ERR_IO, "short atom row");
            dst[k][i] = strtod(t[col[k]].c_str(), nullptr);
        }
    }
    return EDMD_OK;
}

extern "C" int edmd_dump_write(const char* path, int append, double t, int n, double Lx, double Ly,
                               const double* x, const double* y, const double* vx, const double* vy,
                               const double* rad, const double* m, int full_precision) {
    if (!path || n < 0 || !x || !y || !vx || !vy || !rad || !m)
        return edmd_internal_fail(EDMD_ERR_ARG, "bad dump arguments");
    FILE* f = fopen(path, append ? "a" : "w");
    if (!f) return edmd_internal_fail(EDMD_ERR_IO, "cannot open dump for writing");
    if (full_precision) {
        fprintf(f,
                "ITEM: TIMESTEP\n%.17g\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n"
                "0 %.17g\n0 %.17g\n0 0\nITEM: ATOMS id type x y vx vy radius m coll\n",
                t, n, Lx, Ly);
        for (int i = 0; i < n; i++)
            fprintf(f, "%d 1 %.17g %.17g %.17g %.17g %.17g %.17g 1\n", i, x[i], y[i], vx[i], vy[i],
                    rad[i], m[i]);
    } else {
        /* saveTXT, EDMD.c:5036, 5124 */
        fprintf(f,
                "ITEM: TIMESTEP\n%lf\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n"
                "0 %lf\n0 %lf\n0 0\nITEM: ATOMS id type x y vx vy radius m coll\n",
                t, n, Lx, Ly);
        for (int i = 0; i < n; i++)
            fprintf(f, "%d %d %.3lf %lf %lf %lf %lf %lf %d\n", i, 1, x[i], y[i], vx[i], vy[i], rad[i],
                    m[i], 1);
    }
    int bad = ferror(f);
    fclose(f);
    return bad ? edmd_internal_fail(EDMD_ERR_IO, "write error") : EDMD_OK;
}
