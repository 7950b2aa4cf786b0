// Writes a small snapshot file through the driver's run-time HDF5 writer
// (driver/hdf5.h): fields A[i][j] = i + j/1000, B = -A, C = i*j, 3 x 5 cells.
// Exit 0 and prints the library used; exit 2 when no HDF5 library is found.
#include <cstdio>
#include <vector>

#include "hdf5.h"

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    const cfsdrv::Hdf5* h = cfsdrv::Hdf5::get();
    if (!h) {
        std::printf("no hdf5\n");
        return 2;
    }
    const int nx = 3, ny = 5;
    std::vector<double> a(nx * ny), b(nx * ny), c(nx * ny);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            a[i * ny + j] = i + j / 1000.0;
            b[i * ny + j] = -a[i * ny + j];
            c[i * ny + j] = i * j;
        }
    const bool ok = h->write(argv[1], nx, ny, {{"A", a.data()}, {"B", b.data()}, {"C", c.data()}});
    std::printf("%s %s\n", ok ? "ok" : "failed", h->path().c_str());
    return ok ? 0 : 1;
}
