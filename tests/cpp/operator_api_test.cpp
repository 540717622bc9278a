// Drives the reference-shaped C++ API (include/powergrid/powergrid.hpp) the
// way PowerGridIsmrmrd.cpp:242-251 does: Gnufft -> TimeSegmentation -> SENSE
// -> reconSolve.  Inputs come from a binary file written by
// tests/test_cpp_api.py; outputs go to a second file that the test compares
// against the oracle.
//
// input : int32 N1(=Nx=Ny), M, nc, L, niter; float beta;
//         float kx[M], ky[M], t[M], fm[N]; complex<float> sens[N*nc], x[N], y[M*nc]
// output: complex<float> fwd[M*nc], adj[N], recon[N]; int32 iterations;
//         then the direct-reconstruction calls (directRecon.cpp): float
//         dcf[M]; complex<float> coil images[N*nc]; float sos[N]; and
//         G's spatial pair on coil 0's data: complex<float> grid[Gx*Gy],
//         samples[M]; and <out.bin>_img_mag.nii, <out.bin>_img_phs.nii
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "powergrid/powergrid.hpp"

using namespace powergrid;

template <class T>
static void rd(std::ifstream& f, T* p, size_t n)
{
    f.read(reinterpret_cast<char*>(p), sizeof(T) * n);
    if (!f) throw std::runtime_error("short input file");
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::cerr << "usage: operator_api_test in.bin out.bin\n";
        return 2;
    }
    try {
        std::ifstream f(argv[1], std::ios::binary);
        int hdr[5];
        float beta;
        rd(f, hdr, 5);
        rd(f, &beta, 1);
        const uword Nx = hdr[0], Ny = hdr[0], Nz = 1, M = hdr[1], nc = hdr[2], L = hdr[3];
        const uword niter = hdr[4], N = Nx * Ny;
        Col<float> kx(M), ky(M), kz(M, 0.0f), t(M), fm(N);
        Col<complex<float>> sens(N * nc), x(N), y(M * nc);
        rd(f, kx.memptr(), M);
        rd(f, ky.memptr(), M);
        rd(f, t.memptr(), M);
        rd(f, fm.memptr(), N);
        rd(f, sens.memptr(), N * nc);
        rd(f, x.memptr(), N);
        rd(f, y.memptr(), M * nc);

        Col<float> ix, iy, iz;
        initImageSpaceCoords(ix, iy, iz, Nx, Ny, Nz);
        Gnufft<float> G(kx.n_rows, (float)2.0, Nx, Ny, Nz, kx, ky, kz, ix, iy, iz);
        TimeSegmentation<float, Gnufft<float>> A(G, fm, t, kx.n_rows, Nx * Ny * Nz, L, 1, 1);
        SENSE<float, TimeSegmentation<float, Gnufft<float>>> Sg(A, sens, kx.n_rows, Nx * Ny * Nz, nc);
        QuadPenalty<float> R(Nx, Ny, Nz, beta, 2);

        Col<complex<float>> fwd = Sg * x;
        Col<complex<float>> adj = Sg / y;
        Col<complex<float>> img = reconSolve<float, SENSE<float, TimeSegmentation<float, Gnufft<float>>>,
                                             QuadPenalty<float>>(y, Sg, R, kx, ky, kz, Nx, Ny, Nz, t,
                                                                 niter);
        pg_cg_info info{};
        Col<float> W;
        Col<complex<float>> x0(N, complex<float>(0, 0));
        solve_pwls_pcg<float>(x0, Sg, W, y, R, niter, &info);

        std::ofstream o(argv[2], std::ios::binary);
        o.write(reinterpret_cast<const char*>(fwd.memptr()), sizeof(complex<float>) * fwd.size());
        o.write(reinterpret_cast<const char*>(adj.memptr()), sizeof(complex<float>) * adj.size());
        o.write(reinterpret_cast<const char*>(img.memptr()), sizeof(complex<float>) * img.size());
        const int it = info.iterations;
        o.write(reinterpret_cast<const char*>(&it), sizeof(int));

        Col<float> dcf = calc3DDensityCompensation<float>(kx, ky, kz, Nx, Nz);
        Col<complex<float>> coil = gridCoilImages<float>(G, dcf, y, nc);
        Col<float> sos = calcSumOfSquaresImage<float>(coil, Nx, Nz, 1, nc);
        Col<complex<float>> y0(M);
        for (uword i = 0; i < M; ++i) y0(i) = y(i);
        Col<complex<float>> grid = G.adjointSpatialInterp(y0);
        Col<complex<float>> samp = G.forwardSpatialInterp(grid);
        o.write(reinterpret_cast<const char*>(dcf.memptr()), sizeof(float) * dcf.size());
        o.write(reinterpret_cast<const char*>(coil.memptr()), sizeof(complex<float>) * coil.size());
        o.write(reinterpret_cast<const char*>(sos.memptr()), sizeof(float) * sos.size());
        o.write(reinterpret_cast<const char*>(grid.memptr()), sizeof(complex<float>) * grid.size());
        o.write(reinterpret_cast<const char*>(samp.memptr()), sizeof(complex<float>) * samp.size());
        // the reconstruction as magnitude/phase NIfTI files beside out.bin
        writeNiftiMagPhsImage<float>(std::string(argv[2]) + "_img", img, Nx, Ny, Nz);
        std::cout << "operator_api_test ok: " << it << " iterations\n";
    } catch (const std::exception& e) {
        std::cerr << "operator_api_test failed: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
