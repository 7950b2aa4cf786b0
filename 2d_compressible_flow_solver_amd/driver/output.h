// output.h -- snapshots in the reference's field layout (Output::WriteToHDF5,
// Output.cpp:32-88): datasets Mu, Mv, Rho, Energy, Pressure, FluxMu, FluxMv,
// FluxRho, FluxEnergy, Udir, Vdir, Tdir, each [Nx][Ny] x-major FP64.  One
// directory per snapshot holds them as snapshot.h5 (serial HDF5, the
// reference's dataset names in the root group) when an HDF5 C library is
// found at run time (hdf5.h), and always as one NumPy .npy file per dataset (a
// short text header, then the raw little-endian doubles), with meta.json
// naming the stage, time, sizes and the HDF5 file if one was written.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cfs.h"
#include "hdf5.h"

namespace cfsdrv {

inline bool write_npy(const std::string& path, const double* data, int nx, int ny) {
    std::string hdr = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + std::to_string(nx) + ", " +
                      std::to_string(ny) + "), }";
    // magic (6) + version (2) + header length (2) + header, padded with
    // spaces and a newline to a multiple of 64 bytes
    size_t total = 10 + hdr.size() + 1;
    hdr.append((64 - total % 64) % 64, ' ');
    hdr.push_back('\n');
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    const uint16_t len = (uint16_t)hdr.size();
    const unsigned char lenb[2] = {(unsigned char)(len & 0xff), (unsigned char)(len >> 8)};
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(lenb, 1, 2, f) == 2 &&
              std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
              std::fwrite(data, sizeof(double), (size_t)nx * ny, f) == (size_t)nx * ny;
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

class Output {
public:
    // Writes the current state of h, the RHS its last stage applied and the
    // pressure into dir (created by the caller).  Cells that are not fluid
    // carry RHS 0; solid cells carry pressure 0 (never computed in the
    // reference); gradients are 0 (the reference never computes them).
    static int write(cfs_solver* h, const cfs_config& cfg, const std::vector<uint8_t>& cell_type,
                     const std::string& dir, int stage, double t) {
        const size_t n = (size_t)cfg.nx * cfg.ny;
        std::vector<double> U(4 * n), R(4 * n), P(n, 0.0), Z(n, 0.0);
        int rc = cfs_get_state(h, U.data());
        if (rc) return rc;
        rc = cfs_last_rhs(h, R.data());
        if (rc) return rc;
        for (size_t c = 0; c < n; ++c) {
            if (!cell_type.empty() && cell_type[c] == CFS_CELL_SOLID) continue;
            // EulerPhase::computePressurefromOther (EulerPhase.cpp:130-133)
            P[c] = (cfg.gamma - 1) *
                   (U[3 * n + c] - 0.5 * (U[n + c] * U[n + c] + U[2 * n + c] * U[2 * n + c]) / U[c]);
        }
        struct Item { const char* name; const double* p; };
        const Item items[] = {{"Mu", &U[n]},         {"Mv", &U[2 * n]},     {"Rho", &U[0]},
                              {"Energy", &U[3 * n]}, {"Pressure", P.data()}, {"FluxMu", &R[n]},
                              {"FluxMv", &R[2 * n]}, {"FluxRho", &R[0]},     {"FluxEnergy", &R[3 * n]},
                              {"Udir", Z.data()},    {"Vdir", Z.data()},     {"Tdir", Z.data()}};
        for (const Item& it : items)
            if (!write_npy(dir + "/" + it.name + ".npy", it.p, cfg.nx, cfg.ny)) return CFS_E_ARG;
        std::string h5 = "null";
        if (const Hdf5* lib = Hdf5::get()) {
            std::vector<Hdf5::Field> fields;
            for (const Item& it : items) fields.push_back(Hdf5::Field{it.name, it.p});
            if (!lib->write(dir + "/snapshot.h5", cfg.nx, cfg.ny, fields)) return CFS_E_ARG;
            h5 = "\"snapshot.h5\"";
        }
        FILE* f = std::fopen((dir + "/meta.json").c_str(), "w");
        if (!f) return CFS_E_ARG;
        std::fprintf(f,
                     "{\"stage\": %d, \"t\": %.17g, \"nx\": %d, \"ny\": %d, \"layout\": \"[nx][ny] x-major float64\", "
                     "\"datasets\": [\"Mu\", \"Mv\", \"Rho\", \"Energy\", \"Pressure\", \"FluxMu\", \"FluxMv\", "
                     "\"FluxRho\", \"FluxEnergy\", \"Udir\", \"Vdir\", \"Tdir\"], \"hdf5\": %s}\n",
                     stage, t, cfg.nx, cfg.ny, h5.c_str());
        std::fclose(f);
        return CFS_OK;
    }
};

}  // namespace cfsdrv
