// og_io.cpp -- the reference's on-disk outputs either side of the MD path:
// the restart file (WriteRestart::write_restart, WriteTrr.cpp:258-313), the
// Presto trajectory (WriteTrrPresto::write_trr, WriteTrr.cpp:179-254) and the
// GROMACS .trr trajectory (WriteTrrGromacs::write_trr, WriteTrr.cpp:33-172).
// Byte layouts as the reference writes them on x86-64 (little endian, Fortran
// style int record markers); values come in as the engine holds them.
#include "../../include/og_engine.h"

#include <cstdio>
#include <cstring>
#include <vector>

void og_set_error(const char *msg);  // og_engine.cpp

namespace {

struct Out {
    std::vector<char> b;
    template <class T> void put(const T &v) {
        const char *p = (const char *)&v;
        b.insert(b.end(), p, p + sizeof(T));
    }
    void put_bytes(const void *p, size_t n) { b.insert(b.end(), (const char *)p, (const char *)p + n); }
};

og_status flush(const char *path, const Out &o, const char *mode) {
    FILE *f = std::fopen(path, mode);
    if (!f) {
        char m[320];
        std::snprintf(m, sizeof m, "cannot open %s for writing", path);
        og_set_error(m);
        return OG_E_ARG;
    }
    const size_t w = std::fwrite(o.b.data(), 1, o.b.size(), f);
    const int    c = std::fclose(f);
    if (w != o.b.size() || c != 0) {
        og_set_error("short write");
        return OG_E_ARG;
    }
    return OG_OK;
}

}  // namespace

extern "C" {

og_status og_write_restart(const char *path, const char *title, int n_atoms, int n_steps, double time,
                           double e_potential, double e_kinetic, const double *crd, const double *vel) {
    if (!path || n_atoms <= 0 || !crd || !vel) return OG_E_ARG;
    Out o;
    // title record: 80 bytes (the reference copies its version string into an
    // uninitialised buffer; here the rest is zero)
    char t[80];
    std::memset(t, 0, sizeof t);
    if (title) std::strncpy(t, title, sizeof t - 1);
    o.put<int>(80);
    o.put_bytes(t, 80);
    o.put<int>(80);
    // #atoms, #velocities
    o.put<int>(8);
    o.put<int>(n_atoms);
    o.put<int>(n_atoms);
    o.put<int>(8);
    // step, time, total, kinetic, potential
    o.put<int>(36);
    o.put<int>(n_steps);
    o.put<double>(time);
    o.put<double>(e_kinetic + e_potential);
    o.put<double>(e_kinetic);
    o.put<double>(e_potential);
    o.put<int>(36);
    const int rec = n_atoms * 3 * 8;
    o.put<int>(rec);
    o.put_bytes(crd, sizeof(double) * 3 * (size_t)n_atoms);
    o.put<int>(rec);
    o.put<int>(rec);
    o.put_bytes(vel, sizeof(double) * 3 * (size_t)n_atoms);
    o.put<int>(rec);
    return flush(path, o, "wb");
}

og_status og_write_trr_presto_frame(const char *path, int append, int n_atoms, int step, float time,
                                    float cpu_time, float total_e, float kinetic_e, float temperature,
                                    float potential_e, float vdw_e, const double *crd, int n_group,
                                    const int *group) {
    if (!path || n_atoms <= 0 || !crd || (n_group > 0 && !group)) return OG_E_ARG;
    Out o;
    // header record: step, time, cpu time, total, kinetic, temperature,
    // potential, rmsf (0), vdw, hyd (0), rmsd (0); the reference closes it
    // with a 0 marker, not 44
    o.put<int>(44);
    o.put<int>(step);
    o.put<float>(time);
    o.put<float>(cpu_time);
    o.put<float>(total_e);
    o.put<float>(kinetic_e);
    o.put<float>(temperature);
    o.put<float>(potential_e);
    o.put<float>(0.0f);
    o.put<float>(vdw_e);
    o.put<float>(0.0f);
    o.put<float>(0.0f);
    o.put<int>(0);
    const int n   = n_group > 0 ? n_group : n_atoms;
    const int rec = n * 3 * 4;
    o.put<int>(rec);
    for (int k = 0; k < n; k++) {
        const int i = n_group > 0 ? group[k] : k;
        if (i < 0 || i >= n_atoms) return OG_E_ARG;
        for (int d = 0; d < 3; d++) o.put<float>((float)crd[3 * (size_t)i + d]);
    }
    o.put<int>(rec);
    return flush(path, o, append ? "ab" : "wb");
}

og_status og_write_trr_gromacs_frame(const char *path, int append, int n_atoms, int step, double time, double lx,
                                     double ly, double lz, const double *crd, const double *vel, const double *force,
                                     int n_group, const int *group) {
    if (!path || n_atoms <= 0 || (!crd && !vel && !force) || (n_group > 0 && !group)) return OG_E_ARG;
    // the reference sizes every block by n_atoms (also when it writes a group)
    const int box_size = 9 * (int)sizeof(double);
    const int x_size   = crd ? n_atoms * 3 * (int)sizeof(double) : 0;
    const int v_size   = vel ? n_atoms * 3 * (int)sizeof(double) : 0;
    const int f_size   = force ? n_atoms * 3 * (int)sizeof(double) : 0;
    Out       o;
    o.put<int>(1993);
    o.put<int>(13);
    o.put<int>(12);
    o.put_bytes("GMX_trn_file", 12);
    o.put<int>(0);
    o.put<int>(0);
    o.put<int>(box_size);
    for (int k = 0; k < 4; k++) o.put<int>(0);
    o.put<int>(x_size);
    o.put<int>(v_size);
    o.put<int>(f_size);
    o.put<int>(n_atoms);
    o.put<int>(step);
    o.put<int>(0);
    o.put<double>(time);
    o.put<double>(0.0);
    // box in nm
    const double bx[9] = {lx * 0.1, 0, 0, 0, ly * 0.1, 0, 0, 0, lz * 0.1};
    for (double v : bx) o.put<double>(v);
    const int n = n_group > 0 ? n_group : n_atoms;
    for (const double *a : {crd, vel, force}) {
        if (!a) continue;
        for (int k = 0; k < n; k++) {
            const int i = n_group > 0 ? group[k] : k;
            if (i < 0 || i >= n_atoms) return OG_E_ARG;
            for (int d = 0; d < 3; d++) o.put<double>(a[3 * (size_t)i + d] * 0.1);
        }
    }
    return flush(path, o, append ? "ab" : "wb");
}

}  // extern "C"
