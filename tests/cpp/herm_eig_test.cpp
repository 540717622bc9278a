// Reads n and a row-major n x n complex Hermitian matrix (binary doubles,
// re/im interleaved) from argv[1]; writes the eigenvalues and the row-major
// eigenvector matrix of powergrid_amd/csrc/pg_herm.h to argv[2].
#include <complex>
#include <cstdio>
#include <fstream>
#include <vector>

#include "../../powergrid_amd/csrc/pg_herm.h"

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    std::ifstream f(argv[1], std::ios::binary);
    int n = 0;
    f.read(reinterpret_cast<char*>(&n), sizeof n);
    std::vector<std::complex<double>> A((size_t)n * n), V;
    f.read(reinterpret_cast<char*>(A.data()), sizeof(std::complex<double>) * A.size());
    if (!f) return 1;
    std::vector<double> w;
    pg::herm_eig(n, A, w, V);
    std::ofstream o(argv[2], std::ios::binary);
    o.write(reinterpret_cast<const char*>(w.data()), sizeof(double) * n);
    o.write(reinterpret_cast<const char*>(V.data()), sizeof(std::complex<double>) * V.size());
    return 0;
}
