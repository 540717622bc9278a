/*
 * pg_oracle_impl.h -- body of the CPU restatement, instantiated twice by
 * pg_oracle.c (R = float, SUF = f and R = double, SUF = d), the way the
 * reference instantiates its templates for float and double.
 *
 * TEST INFRASTRUCTURE ONLY (see pg_oracle.h).
 */

#define FN(name) CAT(name, SUF)

/* ------------------------------------------------------------------ LUT -- */

/* griddingSupport.cpp:33-48: Numerical Recipes I0 polynomial.  The reference
 * calls it with a double argument even for float (the sqrt(1.0 - v) promotes),
 * so bessi0<double> is what runs; it is shared by both instantiations. */

size_t FN(orc_lut_size)(R width) { return (size_t)(10000 * width); }

/* griddingSupport.cpp:52-76 */
void FN(orc_lut)(R beta, R width, R* lut)
{
    size_t S = FN(orc_lut_size)(width);
    for (size_t k = 0; k < S; ++k) {
        R v = (R)k / (R)S;
        double arg = (double)beta * sqrt(1.0 - (double)v);
        lut[k] = (R)(orc_bessi0(arg) / (double)width);
    }
}

/* Gnufft.cpp:85-88 (MRI_PI is a float literal in the reference's default
 * build; PI_R is that literal for float and the double value for double) */
R FN(orc_kb_beta)(R gridos, R width)
{
    return (R)((double)PI_R * sqrt((gridos - 0.5) * (gridos - 0.5) *
                                   (width * width * 4.0) / (gridos * gridos) - 0.8));
}

/* ------------------------------------------------------------------ FFT -- */

/* Mixed-radix Stockham autosort FFT (radices 4, 2, 3 and a plain DFT for
 * any other factor), twiddles from double-precision cos/sin, butterflies in
 * R.  It stands in for FFTW's unnormalised C2C (fftCPU.cpp:48-106) and is
 * pinned to numpy.fft by tests/test_oracle_core.py. */
typedef struct {
    int n, sign, nst;
    int radix[40];
    R* tw[40];   /* per stage: (k*radix + r) -> w^(r*k), interleaved */
    R* root[40]; /* per stage (generic radix only): w_p^(m), m < radix */
} FN(fftplan);

#define FFT_CACHE 64
static FN(fftplan)* FN(plan_cache)[FFT_CACHE];

static FN(fftplan)* FN(fft_plan)(int n, int sign)
{
    FN(fftplan)* p = NULL;
#pragma omp critical(orc_fft_plan)
    {
        for (int i = 0; i < FFT_CACHE && FN(plan_cache)[i]; ++i)
            if (FN(plan_cache)[i]->n == n && FN(plan_cache)[i]->sign == sign) {
                p = FN(plan_cache)[i];
                break;
            }
        if (!p) {
            p = (FN(fftplan)*)calloc(1, sizeof(*p));
            p->n = n;
            p->sign = sign;
            int rem = n, Ns = 1;
            while (rem > 1) {
                int Rdx;
                if (rem % 4 == 0) Rdx = 4;
                else if (rem % 2 == 0) Rdx = 2;
                else if (rem % 3 == 0) Rdx = 3;
                else {
                    Rdx = rem;
                    for (int q = 5; q * q <= rem; q += 2)
                        if (rem % q == 0) { Rdx = q; break; }
                }
                int st = p->nst++;
                p->radix[st] = Rdx;
                p->tw[st] = (R*)malloc(sizeof(R) * 2 * (size_t)Ns * Rdx);
                for (int k = 0; k < Ns; ++k)
                    for (int r = 0; r < Rdx; ++r) {
                        double ang = sign * 2.0 * M_PI * (double)(r * k) / (double)(Ns * Rdx);
                        p->tw[st][2 * (k * Rdx + r)] = (R)cos(ang);
                        p->tw[st][2 * (k * Rdx + r) + 1] = (R)sin(ang);
                    }
                p->root[st] = (R*)malloc(sizeof(R) * 2 * (size_t)Rdx);
                for (int m = 0; m < Rdx; ++m) {
                    double ang = sign * 2.0 * M_PI * (double)m / (double)Rdx;
                    p->root[st][2 * m] = (R)cos(ang);
                    p->root[st][2 * m + 1] = (R)sin(ang);
                }
                Ns *= Rdx;
                rem /= Rdx;
            }
            int slot = 0;
            while (slot < FFT_CACHE && FN(plan_cache)[slot]) ++slot;
            if (slot < FFT_CACHE) FN(plan_cache)[slot] = p;
        }
    }
    return p;
}

/* x: data (length n), y: work buffer (length n, plus 2*radix scratch) */
static void FN(fft_exec)(const FN(fftplan) * p, R* x, R* y)
{
    int n = p->n, Ns = 1, sign = p->sign;
    R* in = x;
    R* out = y;
    R vr[64], vi[64];
    for (int st = 0; st < p->nst; ++st) {
        int Rdx = p->radix[st];
        int stride = n / Rdx;
        const R* tw = p->tw[st];
        for (int j = 0; j < stride; ++j) {
            int k = j % Ns;
            R* pvr = vr;
            R* pvi = vi;
            R* big = NULL;
            if (Rdx > 64) {
                big = (R*)malloc(sizeof(R) * 2 * (size_t)Rdx);
                pvr = big;
                pvi = big + Rdx;
            }
            for (int r = 0; r < Rdx; ++r) {
                R ar = in[2 * (j + r * stride)], ai = in[2 * (j + r * stride) + 1];
                R c = tw[2 * (k * Rdx + r)], s = tw[2 * (k * Rdx + r) + 1];
                pvr[r] = ar * c - ai * s;
                pvi[r] = ar * s + ai * c;
            }
            int base = (j / Ns) * Ns * Rdx + k;
            if (Rdx == 4) {
                R t0r = pvr[0] + pvr[2], t0i = pvi[0] + pvi[2];
                R t1r = pvr[0] - pvr[2], t1i = pvi[0] - pvi[2];
                R t2r = pvr[1] + pvr[3], t2i = pvi[1] + pvi[3];
                R dr = pvr[1] - pvr[3], di = pvi[1] - pvi[3];
                /* t3 = (sign * i) * (a1 - a3) */
                R t3r = -sign * di, t3i = sign * dr;
                out[2 * base] = t0r + t2r;
                out[2 * base + 1] = t0i + t2i;
                out[2 * (base + Ns)] = t1r + t3r;
                out[2 * (base + Ns) + 1] = t1i + t3i;
                out[2 * (base + 2 * Ns)] = t0r - t2r;
                out[2 * (base + 2 * Ns) + 1] = t0i - t2i;
                out[2 * (base + 3 * Ns)] = t1r - t3r;
                out[2 * (base + 3 * Ns) + 1] = t1i - t3i;
            } else if (Rdx == 2) {
                out[2 * base] = pvr[0] + pvr[1];
                out[2 * base + 1] = pvi[0] + pvi[1];
                out[2 * (base + Ns)] = pvr[0] - pvr[1];
                out[2 * (base + Ns) + 1] = pvi[0] - pvi[1];
            } else if (Rdx == 3) {
                const R h = (R)0.86602540378443864676;
                R t1r = pvr[1] + pvr[2], t1i = pvi[1] + pvi[2];
                R t2r = pvr[0] - (R)0.5 * t1r, t2i = pvi[0] - (R)0.5 * t1i;
                R ur = sign * h * (pvr[1] - pvr[2]), ui = sign * h * (pvi[1] - pvi[2]);
                out[2 * base] = pvr[0] + t1r;
                out[2 * base + 1] = pvi[0] + t1i;
                out[2 * (base + Ns)] = t2r - ui;
                out[2 * (base + Ns) + 1] = t2i + ur;
                out[2 * (base + 2 * Ns)] = t2r + ui;
                out[2 * (base + 2 * Ns) + 1] = t2i - ur;
            } else {
                const R* rt = p->root[st];
                for (int q = 0; q < Rdx; ++q) {
                    R sr = 0, si = 0;
                    for (int r = 0; r < Rdx; ++r) {
                        int m = (r * q) % Rdx;
                        sr += pvr[r] * rt[2 * m] - pvi[r] * rt[2 * m + 1];
                        si += pvr[r] * rt[2 * m + 1] + pvi[r] * rt[2 * m];
                    }
                    out[2 * (base + q * Ns)] = sr;
                    out[2 * (base + q * Ns) + 1] = si;
                }
            }
            if (big) free(big);
        }
        Ns *= Rdx;
        R* t = in;
        in = out;
        out = t;
    }
    if (in != x) memcpy(x, in, sizeof(R) * 2 * (size_t)n);
}

void FN(orc_fft1d)(R* data, int n, int sign)
{
    const FN(fftplan)* p = FN(fft_plan)(n, sign);
    R* tmp = (R*)malloc(sizeof(R) * 2 * (size_t)n);
    FN(fft_exec)(p, data, tmp);
    free(tmp);
}

/* transform along one axis of a (n0 fastest, n1, n2) array */
static void FN(fft_axis)(R* data, int n0, int n1, int n2, int axis, int sign)
{
    int n = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
    if (n <= 1) return;
    const FN(fftplan)* p = FN(fft_plan)(n, sign);
    long stride = axis == 0 ? 1 : (axis == 1 ? n0 : (long)n0 * n1);
    long lines = (long)n0 * n1 * n2 / n;
#pragma omp parallel num_threads(orc_nthreads())
    {
        R* buf = (R*)malloc(sizeof(R) * 4 * (size_t)n);
#pragma omp for schedule(static)
        for (long ln = 0; ln < lines; ++ln) {
            long base;
            if (axis == 0) base = ln * n0;
            else if (axis == 1) base = (ln % n0) + (ln / n0) * (long)n0 * n1;
            else base = ln;
            if (axis == 0) {
                FN(fft_exec)(p, data + 2 * base, buf);
            } else {
                for (int i = 0; i < n; ++i) {
                    buf[2 * i] = data[2 * (base + i * stride)];
                    buf[2 * i + 1] = data[2 * (base + i * stride) + 1];
                }
                FN(fft_exec)(p, buf, buf + 2 * n);
                for (int i = 0; i < n; ++i) {
                    data[2 * (base + i * stride)] = buf[2 * i];
                    data[2 * (base + i * stride) + 1] = buf[2 * i + 1];
                }
            }
        }
        free(buf);
    }
}

/* fftCPU.cpp:48-106: fftwf_plan_dft_2d(ny, nx, ...) / _3d(nz, ny, nx, ...):
 * nx is the contiguous dimension as FFTW sees the buffer. */
void FN(orc_fft)(R* data, int nx, int ny, int nz, int sign)
{
    FN(fft_axis)(data, nx, ny, nz, 0, sign);
    FN(fft_axis)(data, nx, ny, nz, 1, sign);
    FN(fft_axis)(data, nx, ny, nz, 2, sign);
}

/* ------------------------------------------------------- grid support -- */

static R FN(kb_ft)(R W, R beta, R u)
{
    /* griddingSupport.cpp:249-260 */
    R e = PI_R * PI_R * W * W * u * u - beta * beta;
    if (e >= 0) return (R)(sin(sqrt(e)) / sqrt(e));
    return (R)(sinh(sqrt((R)-1.0 * e)) / sqrt((R)-1.0 * e));
}

/* griddingSupport.cpp:209-368 */
void FN(orc_deapod)(R* dst, const R* src, int Nx, int Ny, int Nz, R W, R beta, R os)
{
    if (Nz == 1) {
        int n = Nx * Ny;
        for (int i = 0; i < 2 * n; ++i) dst[i] = (R)0.0;
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
        for (int X = 0; X < Nx; X++) {
            for (int Y = 0; Y < Ny; Y++) {
                R gy = (R)(Y - (Ny / 2)) / (R)Ny;
                R gx = (R)(X - (Nx / 2)) / (R)Nx;
                R kern = FN(kb_ft)(W, beta, gx) * FN(kb_ft)(W, beta, gy);
                int ci = Y + X * Ny;
                if (!isnan(kern)) {
                    R os2 = os * os;
                    dst[2 * ci] = src[2 * ci] / kern * ((R)1.0 / os2);
                    dst[2 * ci + 1] = src[2 * ci + 1] / kern * ((R)1.0 / os2);
                } else {
                    dst[2 * ci] = 0;
                    dst[2 * ci + 1] = 0;
                }
            }
        }
    } else {
        long n = (long)Nx * Ny * Nz;
        for (long i = 0; i < 2 * n; ++i) dst[i] = (R)0.0;
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
        for (int Z = 0; Z < Nz; Z++) {
            for (int X = 0; X < Nx; X++) {
                for (int Y = 0; Y < Ny; Y++) {
                    R gz = (R)((Z) - ((R)Nz / 2.0)) / (R)Nz;
                    R gy = (R)((Y) - ((R)Ny / 2.0)) / (R)Ny;
                    R gx = (R)((X) - ((R)Nx / 2.0)) / (R)Nx;
                    R kern = FN(kb_ft)(W, beta, gx) * FN(kb_ft)(W, beta, gy) *
                             FN(kb_ft)(W, beta, gz);
                    long ci = (long)Z * Ny * Nx + (long)X * Ny + Y;
                    if (!isnan(kern)) {
                        R os3 = os * os * os;
                        dst[2 * ci] = (R)(src[2 * ci] / kern * (1.0 / os3));
                        dst[2 * ci + 1] = (R)(src[2 * ci + 1] / kern * (1.0 / os3));
                    } else {
                        dst[2 * ci] = 0;
                        dst[2 * ci + 1] = 0;
                    }
                }
            }
        }
    }
}

/* griddingSupport.cpp:456-547.  The 2-D index is x-fastest in the reference
 * (:488-489); it is restated as written. */
void FN(orc_zero_pad)(R* dst, const R* src, int Nx, int Ny, int Nz, R os)
{
    if (Nz == 1) {
        int offY = (int)(((R)Ny * os / (R)2.0) - ((R)Ny / (R)2.0));
        int offX = (int)(((R)Nx * os / (R)2.0) - ((R)Nx / (R)2.0));
        int dsz = (int)(Nx * os * Ny * os);
        int rowX = (int)(Nx * os);
        for (int j = 0; j < 2 * dsz; ++j) dst[j] = 0;
        for (int dY = 0; dY < Ny; dY++)
            for (int dX = 0; dX < Nx; dX++) {
                int di = (dY + offY) * rowX + (dX + offX);
                int si = dY * Nx + dX;
                dst[2 * di] = src[2 * si];
                dst[2 * di + 1] = src[2 * si + 1];
            }
    } else {
        int offY = (int)(((R)Ny * os / (R)2.0) - ((R)Ny / (R)2.0));
        int offX = (int)(((R)Nx * os / (R)2.0) - ((R)Nx / (R)2.0));
        int offZ = (int)(((R)Nz * os / (R)2.0) - ((R)Nz / (R)2.0));
        long dsz = (long)(Nx * os * Ny * os * Nz * os);
        long pXY = (long)(Nx * Ny * os * os);
        long pY = (long)(Ny * os);
        for (long j = 0; j < 2 * dsz; ++j) dst[j] = 0;
        for (int dZ = 0; dZ < Nz; dZ++)
            for (int dY = 0; dY < Ny; dY++)
                for (int dX = 0; dX < Nx; dX++) {
                    long di = (long)(dZ + offZ) * pXY + (long)(dX + offX) * pY + (dY + offY);
                    long si = (long)dZ * Nx * Ny + (long)dX * Ny + dY;
                    dst[2 * di] = src[2 * si];
                    dst[2 * di + 1] = src[2 * si + 1];
                }
    }
}

/* griddingSupport.cpp:371-453 */
void FN(orc_crop)(R* dst, const R* src, int Nx, int Ny, int Nz, int Gx, int Gy, int Gz)
{
    if (Nz == 1) {
        int offY = (int)(((R)Gy / 2.0f) - ((R)Ny / 2.0f));
        int offX = (int)(((R)Gx / 2.0f) - ((R)Nx / 2.0f));
        for (int dX = 0; dX < Nx; dX++)
            for (int dY = 0; dY < Ny; dY++) {
                int di = dX * Ny + dY;
                int si = (dX + offX) * Gy + (dY + offY);
                dst[2 * di] = src[2 * si];
                dst[2 * di + 1] = src[2 * si + 1];
            }
    } else {
        int offY = (int)(((R)Gy / (R)2.0) - ((R)Ny / (R)2.0));
        int offX = (int)(((R)Gx / (R)2.0) - ((R)Nx / (R)2.0));
        int offZ = (int)(((R)Gz / (R)2.0) - ((R)Nz / (R)2.0));
        for (int dZ = 0; dZ < Nz; dZ++)
            for (int dX = 0; dX < Nx; dX++)
                for (int dY = 0; dY < Ny; dY++) {
                    long di = (long)dZ * Nx * Ny + (long)dX * Ny + dY;
                    long si = (long)(dZ + offZ) * Gx * Gy + (long)(dX + offX) * Gy + (dY + offY);
                    dst[2 * di] = src[2 * si];
                    dst[2 * di + 1] = src[2 * si + 1];
                }
    }
}

/* griddingSupport.cpp:551-593: circshift2 / circshift3 (first dim fastest) */
static void FN(circshift)(R* dst, const R* src, int xd, int yd, int zd, int xs, int ys, int zs)
{
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
    for (int x = 0; x < xd; x++)
        for (int y = 0; y < yd; y++)
            for (int z = 0; z < zd; z++) {
                long ii = (x + xs) % xd, jj = (y + ys) % yd, kk = (z + zs) % zd;
                long di = ii + jj * xd + kk * (long)xd * yd;
                long si = x + (long)y * xd + (long)z * xd * yd;
                dst[2 * di] = src[2 * si];
                dst[2 * di + 1] = src[2 * si + 1];
            }
}

/* griddingSupport.cpp:599-627 (ifftshift3 uses floor for z) */
void FN(orc_fftshift)(R* dst, const R* src, int Gx, int Gy, int Gz)
{
    FN(circshift)(dst, src, Gx, Gy, Gz, (int)floor(Gx / 2.0), (int)floor(Gy / 2.0),
                  (int)floor(Gz / 2.0));
}
void FN(orc_ifftshift)(R* dst, const R* src, int Gx, int Gy, int Gz)
{
    FN(circshift)(dst, src, Gx, Gy, Gz, (int)ceil(Gx / 2.0), (int)ceil(Gy / 2.0),
                  (int)floor(Gz / 2.0));
}

/* ---------------------------------------------------- gridding kernels -- */

static inline R FN(lut_w)(R dist, R W, const R* lut, size_t S)
{
    /* gridding.cpp:80-84: floor-indexed LUT, zero past the end */
    int k0 = (int)((dist * dist * (R)4.0 / (W * W)) * (R)S);
    return (k0 >= (int)S) ? (R)0.0 : lut[k0];
}

static inline void FN(window)(R k, int N, R os, R W, R* shifted, int* lo, int* hi)
{
    /* gridding.cpp:64-73 */
    R s = (os) * (k + ((R)N) / (R)2.0);
    *shifted = s;
    *lo = (int)(fmax((R)0.0, ceil(s - W * (os) / (R)2.0)));
    *hi = (int)(fmin((os * (R)N - (R)1.0), floor(s + W * (os) / (R)2.0)));
}

/* gridding.cpp:217-301 / 304-415 (forward: gather).  Accumulates into
 * samples (the pipeline zeroes it first, gridding.cpp:637-640). */
void FN(orc_grid_forward)(int M, int Nx, int Ny, int Nz, R os, R W, const R* kx,
                          const R* ky, const R* kz, const R* lut, size_t S,
                          const R* grid, R* samples)
{
    int Gy = (int)ceil(os * (R)Ny);
    if (Gy % 2) Gy += 1;
    int Gx = (int)ceil(os * (R)Nx);
    if (Gx % 2) Gx += 1;
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
    for (int i = 0; i < M; i++) {
        R sx, sy, sz = 0;
        int xl, xh, yl, yh, zl = 0, zh = 0;
        FN(window)(kx[i], Nx, os, W, &sx, &xl, &xh);
        FN(window)(ky[i], Ny, os, W, &sy, &yl, &yh);
        if (Nz > 1) FN(window)(kz[i], Nz, os, W, &sz, &zl, &zh);
        R accr = samples[2 * i], acci = samples[2 * i + 1];
        for (int nz = zl; nz <= zh; ++nz) {
            R kbZ = (R)1.0;
            if (Nz > 1) kbZ = FN(lut_w)(fabs(sz - (R)nz) / os, W, lut, S);
            for (int nx = xl; nx <= xh; ++nx) {
                R kbX = FN(lut_w)(fabs(sx - (R)nx) / os, W, lut, S);
                for (int ny = yl; ny <= yh; ++ny) {
                    R kbY = FN(lut_w)(fabs(sy - (R)ny) / os, W, lut, S);
                    R w = (Nz > 1) ? kbX * kbY * kbZ : kbX * kbY;
                    long idx = ny + (long)nx * Gy + (long)nz * Gx * Gy;
                    accr += w * grid[2 * idx];
                    acci += w * grid[2 * idx + 1];
                }
            }
        }
        samples[2 * i] = accr;
        samples[2 * i + 1] = acci;
    }
}

/* gridding.cpp:31-214 (adjoint: spread with atomic updates) */
void FN(orc_grid_adjoint)(int M, int Nx, int Ny, int Nz, R os, R W, const R* kx,
                          const R* ky, const R* kz, const R* lut, size_t S,
                          const R* samples, R* grid)
{
    int Gy = (int)ceil(os * (R)Ny);
    if (Gy % 2) Gy += 1;
    int Gx = (int)ceil(os * (R)Nx);
    if (Gx % 2) Gx += 1;
    /* serial build: plain updates; multicore build: atomic updates */
    const int par = orc_nthreads() > 1;
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
    for (int i = 0; i < M; i++) {
        R sx, sy, sz = 0;
        int xl, xh, yl, yh, zl = 0, zh = 0;
        FN(window)(kx[i], Nx, os, W, &sx, &xl, &xh);
        FN(window)(ky[i], Ny, os, W, &sy, &yl, &yh);
        if (Nz > 1) FN(window)(kz[i], Nz, os, W, &sz, &zl, &zh);
        R pr = samples[2 * i], pi = samples[2 * i + 1];
        for (int nz = zl; nz <= zh; ++nz) {
            R kbZ = (R)1.0;
            if (Nz > 1) kbZ = FN(lut_w)(fabs(sz - (R)nz) / os, W, lut, S);
            for (int nx = xl; nx <= xh; ++nx) {
                R kbX = FN(lut_w)(fabs(sx - (R)nx) / os, W, lut, S);
                for (int ny = yl; ny <= yh; ++ny) {
                    R kbY = FN(lut_w)(fabs(sy - (R)ny) / os, W, lut, S);
                    R w = (Nz > 1) ? kbX * kbY * kbZ : kbX * kbY;
                    long idx = ny + (long)nx * Gy + (long)nz * Gx * Gy;
                    if (par) {
#pragma omp atomic
                        grid[2 * idx] += w * pr;
#pragma omp atomic
                        grid[2 * idx + 1] += w * pi;
                    } else {
                        grid[2 * idx] += w * pr;
                        grid[2 * idx + 1] += w * pi;
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------- Gdft -- */

/* ftCpu.cpp:94-140: phase in R, sin/cos in float */
void FN(orc_ft)(R* out, const R* in, const R* kx, const R* ky, const R* kz,
                const R* ix, const R* iy, const R* iz, const R* fm, const R* t,
                int M, int N)
{
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
    for (int i = 0; i < M; i++) {
        R sumr = 0, sumi = 0;
        R kxtpi = kx[i] * 2 * PI_R;
        R kytpi = ky[i] * 2 * PI_R;
        R kztpi = kz[i] * 2 * PI_R;
        R myti = t[i];
        for (int j = 0; j < N; j++) {
            R expr = (kxtpi * ix[j] + kytpi * iy[j] + kztpi * iz[j] + (fm[j] * myti));
            float se = sinf((float)expr);
            float ce = cosf((float)expr);
            sumr += (ce * in[2 * j]) + (se * in[2 * j + 1]);
            sumi += (-se * in[2 * j]) + (ce * in[2 * j + 1]);
        }
        out[2 * i] = sumr;
        out[2 * i + 1] = sumi;
    }
}

/* ftCpu.cpp:149-215 */
void FN(orc_ift)(R* out, const R* in, const R* kx, const R* ky, const R* kz,
                 const R* ix, const R* iy, const R* iz, const R* fm, const R* t,
                 int M, int N)
{
    R tpi = PI_R * 2.0;
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
    for (int j = 0; j < N; j++) {
        R sumr = 0, sumi = 0;
        R itx = ix[j] * tpi, ity = iy[j] * tpi, itz = iz[j] * tpi;
        R myfmj = fm[j];
        for (int i = 0; i < M; i++) {
            R expr = (kx[i] * itx + ky[i] * ity + kz[i] * itz + (myfmj * t[i]));
            R se = (R)sinf((float)expr);
            R ce = (R)cosf((float)expr);
            sumr += (ce * in[2 * i]) - (se * in[2 * i + 1]);
            sumi += (se * in[2 * i]) + (ce * in[2 * i + 1]);
        }
        out[2 * j] = sumr;
        out[2 * j + 1] = sumi;
    }
}

/* -------------------------------------------------------- operators -- */

typedef struct FN(orc_op_s) {
    int kind; /* 0 nufft, 1 dft, 2 ts, 3 sense, 4 dense, 5 pcsense */
    int n_in, n_out;
    void* ctx;
} FN(orc_op);

typedef struct {
    int M, Nx, Ny, Nz, Gx, Gy, Gz;
    R os, W, beta;
    size_t S;
    R *lut, *kx, *ky, *kz;
    R *img, *img_d, *grid, *grid_d;
} FN(nufft_ctx);

typedef struct {
    int M, N;
    R *kx, *ky, *kz, *ix, *iy, *iz, *fm, *t;
    /* GdftR2 only: field-map difference maps and the image dimensions */
    R *gx, *gy, *gz;
    int numx, numy, numz;
} FN(dft_ctx);

typedef struct {
    FN(orc_op) * G;
    int M, N, L;
    R tau, tmin;
    R* AA;  /* M x L real weights (column-major) */
    R* AAc; /* M x L complex weights (type 2), replacing AA when set */
    R* Wo;  /* N x L complex */
    R* WoH; /* N x L complex */
    R *tmp_in, *tmp_out, *acc;
} FN(ts_ctx);

typedef struct {
    FN(orc_op) * G;
    int M, N, nc;
    R* S; /* N x nc complex */
    R *tmp_in, *tmp_out;
} FN(sense_ctx);

typedef struct {
    int rows, cols;
    R* A;
} FN(dense_ctx);

/* pcSenseTimeSeg.cpp:40-124: Ns shots, each with its own Gnufft and
 * TimeSegmentation over the shot's Nd samples; shot-specific maps
 * SS[c*Ns + s] = S_c exp(-i P_s), CS[c*Ns + s] = conj(S_c) conj(exp(-i P_s)) */
typedef struct {
    int Ns, Nc, Nd, Ni;
    FN(orc_op) * *G;
    FN(orc_op) * *A;
    R* SS;  /* Ni x (Nc*Ns) complex */
    R* CS;
    R *tmp_in, *tmp_out;
} FN(pcsense_ctx);

/* MPI/mpipcSENSETimeSeg.cpp:39-150: the share of one MPI rank -- its tasks
 * (shot, coil), one Gnufft and TimeSegmentation per task -- with the maps
 * S_c exp(-i P_s) formed at each call (:170, :223) */
typedef struct {
    int ntask, Ni, Nd;
    int *shot, *coil;
    FN(orc_op) * *G;
    FN(orc_op) * *A;
    R* S; /* Ni x Nc complex */
    R* P; /* Ni x Ns real */
    R *tmp_in, *tmp_out;
} FN(mpipc_ctx);

static R* FN(dup)(const R* p, size_t n)
{
    R* q = (R*)malloc(sizeof(R) * (n ? n : 1));
    if (p) memcpy(q, p, sizeof(R) * n);
    else memset(q, 0, sizeof(R) * n);
    return q;
}

/* Gnufft.cpp:34-122 (kernelWidth is a parameter here; the reference fixes
 * it to 4.0 at Gnufft.cpp:84) */
void* FN(orc_gnufft_create)(int M, R gridos, int Nx, int Ny, int Nz, const R* kx,
                            const R* ky, const R* kz, R kernelWidth)
{
    FN(nufft_ctx)* c = (FN(nufft_ctx)*)calloc(1, sizeof(*c));
    c->M = M;
    c->Nx = Nx;
    c->Ny = Ny;
    c->Nz = Nz;
    c->os = gridos;
    c->Gx = (int)ceil(gridos * (R)Nx);
    if (c->Gx % 2) c->Gx += 1;
    c->Gy = (int)ceil(gridos * (R)Ny);
    if (c->Gy % 2) c->Gy += 1;
    c->Gz = (Nz == 1) ? Nz : (int)ceil(gridos * (R)Nz);
    c->W = kernelWidth;
    c->beta = FN(orc_kb_beta)(gridos, kernelWidth);
    c->S = FN(orc_lut_size)(kernelWidth);
    c->lut = (R*)malloc(sizeof(R) * c->S);
    FN(orc_lut)(c->beta, kernelWidth, c->lut);
    c->kx = FN(dup)(kx, M);
    c->ky = FN(dup)(ky, M);
    c->kz = FN(dup)(kz, M);
    long n1 = (long)Nx * Ny * Nz;
    long g = (long)c->Gx * c->Gy * c->Gz;
    c->img = FN(dup)(NULL, 2 * n1);
    c->img_d = FN(dup)(NULL, 2 * n1);
    c->grid = FN(dup)(NULL, 2 * g);
    c->grid_d = FN(dup)(NULL, 2 * g);
    FN(orc_op)* op = (FN(orc_op)*)calloc(1, sizeof(*op));
    op->kind = 0;
    op->n_in = (int)n1;
    op->n_out = M;
    op->ctx = c;
    return op;
}

/* gridding.cpp:584-727 (computeFd_CPU_Grid) */
static void FN(nufft_fwd)(FN(nufft_ctx) * c, const R* in, R* out)
{
    long n1 = (long)c->Nx * c->Ny * c->Nz;
    memcpy(c->img, in, sizeof(R) * 2 * n1);
    memset(out, 0, sizeof(R) * 2 * c->M);
    FN(orc_deapod)(c->img_d, c->img, c->Nx, c->Ny, c->Nz, c->W, c->beta, c->os);
    FN(orc_zero_pad)(c->grid, c->img_d, c->Nx, c->Ny, c->Nz, c->os);
    FN(orc_fftshift)(c->grid_d, c->grid, c->Gx, c->Gy, c->Gz);
    if (c->Nz == 1) FN(orc_fft)(c->grid_d, c->Gx, c->Gy, 1, -1);
    else FN(orc_fft)(c->grid_d, c->Gx, c->Gy, c->Gz, -1);
    FN(orc_ifftshift)(c->grid, c->grid_d, c->Gx, c->Gy, c->Gz);
    FN(orc_grid_forward)(c->M, c->Nx, c->Ny, c->Nz, c->os, c->W, c->kx, c->ky, c->kz,
                         c->lut, c->S, c->grid, out);
}

/* gridding.cpp:418-581 (computeFH_CPU_Grid) */
static void FN(nufft_adj)(FN(nufft_ctx) * c, const R* in, R* out)
{
    long g = (long)c->Gx * c->Gy * c->Gz;
    memset(c->grid, 0, sizeof(R) * 2 * g);
    FN(orc_grid_adjoint)(c->M, c->Nx, c->Ny, c->Nz, c->os, c->W, c->kx, c->ky, c->kz,
                         c->lut, c->S, in, c->grid);
    FN(orc_ifftshift)(c->grid_d, c->grid, c->Gx, c->Gy, c->Gz);
    if (c->Nz == 1) FN(orc_fft)(c->grid_d, c->Gx, c->Gy, 1, +1);
    else FN(orc_fft)(c->grid_d, c->Gx, c->Gy, c->Gz, +1);
    FN(orc_fftshift)(c->grid, c->grid_d, c->Gx, c->Gy, c->Gz);
    FN(orc_crop)(c->img_d, c->grid, c->Nx, c->Ny, c->Nz, c->Gx, c->Gy, c->Gz);
    FN(orc_deapod)(out, c->img_d, c->Nx, c->Ny, c->Nz, c->W, c->beta, c->os);
}

/* ftCpuWithGrads.cpp:86-96: sin(pi x)/(pi x), 1 below 1e-4, in R with the
 * float literal pi of the float build */
static inline R FN(sinc_pg)(R x)
{
    if (fabs(x) < 0.0001) return (R)1.0;
    return (R)(sin(PI_R * x) / (PI_R * x));
}

/* GdftR2.cpp:60-107, Cd(FM, dim): d[i + o] - d[i - o] for o <= i < n - o,
 * d[i + o] - d[i] below, d[i] - d[i - o] above, o = 1 (dim 0), Ny (dim 1),
 * Nx*Ny (dim 2), in R.  With Nz == 1 the dim-2 offset is the whole map and
 * the reference indexes outside it (undefined); the map is 0 there. */
void FN(orc_gdftr2_grad)(const R* fm, int nx, int ny, int nz, int dim, R* out)
{
    const long n = (long)nx * ny * nz;
    const long o = dim == 0 ? 1 : dim == 1 ? ny : (long)nx * ny;
    if (o >= n) {
        for (long i = 0; i < n; ++i) out[i] = 0;
        return;
    }
    for (long i = o; i < n - o; ++i) out[i] = fm[i + o] - fm[i - o];
    for (long i = 0; i < o; ++i) out[i] = fm[i + o] - fm[i];
    for (long i = n - o; i < n; ++i) out[i] = fm[i] - fm[i - o];
}

/* ftCpuWithGrads.cpp:98-137: ftCpu with the factor
 * sinc(kx/Nx + Gx t) sinc(ky/Ny + Gy t) sinc(kz/Nz + Gz t) per pair */
static void FN(ft_grads)(const FN(dft_ctx) * c, const R* in, R* out)
{
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
    for (int i = 0; i < c->M; i++) {
        R sumr = 0, sumi = 0;
        R kxtpi = c->kx[i] * 2 * PI_R;
        R kytpi = c->ky[i] * 2 * PI_R;
        R kztpi = c->kz[i] * 2 * PI_R;
        R myti = c->t[i];
        for (int j = 0; j < c->N; j++) {
            R expr = (kxtpi * c->ix[j] + kytpi * c->iy[j] + kztpi * c->iz[j] + (c->fm[j] * myti));
            float se = sinf((float)expr);
            float ce = cosf((float)expr);
            R b = FN(sinc_pg)(c->kx[i] / c->numx + c->gx[j] * myti) *
                  FN(sinc_pg)(c->ky[i] / c->numy + c->gy[j] * myti) *
                  FN(sinc_pg)(c->kz[i] / c->numz + c->gz[j] * myti);
            sumr += b * ((ce * in[2 * j]) + (se * in[2 * j + 1]));
            sumi += b * ((-se * in[2 * j]) + (ce * in[2 * j + 1]));
        }
        out[2 * i] = sumr;
        out[2 * i + 1] = sumi;
    }
}

/* ftCpuWithGrads.cpp:150-212 */
static void FN(ift_grads)(const FN(dft_ctx) * c, const R* in, R* out)
{
    R tpi = PI_R * 2.0;
#pragma omp parallel for num_threads(orc_nthreads()) schedule(static)
    for (int j = 0; j < c->N; j++) {
        R sumr = 0, sumi = 0;
        R itx = c->ix[j] * tpi, ity = c->iy[j] * tpi, itz = c->iz[j] * tpi;
        R myfmj = c->fm[j];
        for (int i = 0; i < c->M; i++) {
            R expr = (c->kx[i] * itx + c->ky[i] * ity + c->kz[i] * itz + (myfmj * c->t[i]));
            R se = (R)sinf((float)expr);
            R ce = (R)cosf((float)expr);
            R b = FN(sinc_pg)(c->kx[i] / c->numx + c->gx[j] * c->t[i]) *
                  FN(sinc_pg)(c->ky[i] / c->numy + c->gy[j] * c->t[i]) *
                  FN(sinc_pg)(c->kz[i] / c->numz + c->gz[j] * c->t[i]);
            sumr += b * ((ce * in[2 * i]) - (se * in[2 * i + 1]));
            sumi += b * ((se * in[2 * i]) + (ce * in[2 * i + 1]));
        }
        out[2 * j] = sumr;
        out[2 * j + 1] = sumi;
    }
}

/* Gdft.cpp:32-49 */
void* FN(orc_gdft_create)(int M, int N, const R* kx, const R* ky, const R* kz,
                          const R* ix, const R* iy, const R* iz, const R* fm, const R* t)
{
    FN(dft_ctx)* c = (FN(dft_ctx)*)calloc(1, sizeof(*c));
    c->M = M;
    c->N = N;
    c->kx = FN(dup)(kx, M);
    c->ky = FN(dup)(ky, M);
    c->kz = FN(dup)(kz, M);
    c->t = FN(dup)(t, M);
    c->ix = FN(dup)(ix, N);
    c->iy = FN(dup)(iy, N);
    c->iz = FN(dup)(iz, N);
    c->fm = FN(dup)(fm, N);
    FN(orc_op)* op = (FN(orc_op)*)calloc(1, sizeof(*op));
    op->kind = 1;
    op->n_in = N;
    op->n_out = M;
    op->ctx = c;
    return op;
}

/* GdftR2.cpp:32-57 */
void* FN(orc_gdftr2_create)(int M, int N, const R* kx, const R* ky, const R* kz, const R* ix,
                            const R* iy, const R* iz, const R* fm, const R* t, int nx, int ny,
                            int nz)
{
    FN(orc_op)* op = (FN(orc_op)*)FN(orc_gdft_create)(M, N, kx, ky, kz, ix, iy, iz, fm, t);
    FN(dft_ctx)* c = (FN(dft_ctx)*)op->ctx;
    c->numx = nx;
    c->numy = ny;
    c->numz = nz;
    c->gx = FN(dup)(NULL, (size_t)N);
    c->gy = FN(dup)(NULL, (size_t)N);
    c->gz = FN(dup)(NULL, (size_t)N);
    FN(orc_gdftr2_grad)(fm, nx, ny, nz, 0, c->gx);
    FN(orc_gdftr2_grad)(fm, nx, ny, nz, 1, c->gy);
    FN(orc_gdftr2_grad)(fm, nx, ny, nz, 2, c->gz);
    return op;
}

/* TimeSegmentation.cpp:36-93, 213-218 (Hanning interpolator, type 1) */
void* FN(orc_ts_create)(void* Gop, const R* fm, const R* tvec, int M, int N, int L,
                        int interptype, int shots)
{
    (void)interptype; /* only the Hanning interpolator (type 1) is restated */
    FN(ts_ctx)* c = (FN(ts_ctx)*)calloc(1, sizeof(*c));
    c->G = (FN(orc_op)*)Gop;
    c->M = M;
    c->N = N;
    c->L = (L == 0 || L == 1) ? 1 : L;
    if (shots < 1) shots = 1;
    R* tv = FN(dup)(tvec, M);
    R tmin = tv[0], tmax = tv[0];
    for (int j = 1; j < M; ++j) {
        if (tv[j] < tmin) tmin = tv[j];
        if (tv[j] > tmax) tmax = tv[j];
    }
    R rangt = tmax - tmin;
    c->tmin = tmin;
    c->tau = (R)(((double)rangt + 2.220446049250313e-16) / (double)c->L);
    for (int j = 0; j < M; ++j) tv[j] = tv[j] - tmin;
    int Lc = c->L;
    c->AA = FN(dup)(NULL, (size_t)M * Lc);
    c->Wo = FN(dup)(NULL, (size_t)2 * N * Lc);
    c->WoH = FN(dup)(NULL, (size_t)2 * N * Lc);
    if (c->L <= 1) {
        c->tau = 0;
        for (int j = 0; j < M; ++j) c->AA[j] = 1;
        for (int n = 0; n < N; ++n) {
            c->Wo[2 * n] = 1;
            c->WoH[2 * n] = 1;
        }
    } else {
        int none = M / shots;
        for (int l = 0; l < Lc; ++l)
            for (int j = 0; j < none; ++j) {
                R d = tv[j] - ((R)(unsigned)l * c->tau);
                R a;
                if (fabs(d) <= c->tau)
                    a = (R)(0.5 + 0.5 * cos(3.141592653589793 * (double)d / (double)c->tau));
                else a = 0;
                for (int s = 0; s < shots; ++s) c->AA[(size_t)l * M + j + (size_t)s * none] = a;
            }
        for (int l = 0; l < Lc; ++l) {
            R sl = (R)(unsigned)l * c->tau;
            sl = sl + c->tmin;
            for (int n = 0; n < N; ++n) {
                R ph = (-fm[n]) * sl;
                c->Wo[2 * ((size_t)l * N + n)] = (R)cos(ph);
                c->Wo[2 * ((size_t)l * N + n) + 1] = (R)sin(ph);
                R phh = fm[n] * sl;
                c->WoH[2 * ((size_t)l * N + n)] = (R)cos(phh);
                c->WoH[2 * ((size_t)l * N + n) + 1] = (R)sin(phh);
            }
        }
    }
    free(tv);
    c->tmp_in = FN(dup)(NULL, 2 * (size_t)(N > M ? N : M));
    c->tmp_out = FN(dup)(NULL, 2 * (size_t)(N > M ? N : M));
    c->acc = FN(dup)(NULL, 2 * (size_t)(N > M ? N : M));
    FN(orc_op)* op = (FN(orc_op)*)calloc(1, sizeof(*op));
    op->kind = 2;
    op->n_in = N;
    op->n_out = M;
    op->ctx = c;
    return op;
}

/* complex interpolator weights (M x L, column-major) replacing the real
 * ones: the min-max interpolators (TimeSegmentation.cpp:94-211) */
int FN(orc_ts_set_aa_complex)(void* tsop, const R* AAc)
{
    FN(orc_op)* op = (FN(orc_op)*)tsop;
    if (op->kind != 2) return -1;
    FN(ts_ctx)* c = (FN(ts_ctx)*)op->ctx;
    free(c->AAc);
    c->AAc = FN(dup)(AAc, 2 * (size_t)c->M * c->L);
    return 0;
}

int FN(orc_ts_tables)(void* tsop, R* AA, R* Wo, R* tau_tmin)
{
    FN(orc_op)* op = (FN(orc_op)*)tsop;
    if (op->kind != 2) return -1;
    FN(ts_ctx)* c = (FN(ts_ctx)*)op->ctx;
    if (AA) memcpy(AA, c->AA, sizeof(R) * (size_t)c->M * c->L);
    if (Wo) memcpy(Wo, c->Wo, sizeof(R) * 2 * (size_t)c->N * c->L);
    if (tau_tmin) {
        tau_tmin[0] = c->tau;
        tau_tmin[1] = c->tmin;
    }
    return c->L;
}

/* SENSE.cpp:32-44 */
void* FN(orc_sense_create)(void* Gop, const R* sens, int M, int N, int nc)
{
    FN(sense_ctx)* c = (FN(sense_ctx)*)calloc(1, sizeof(*c));
    c->G = (FN(orc_op)*)Gop;
    c->M = M;
    c->N = N;
    c->nc = nc;
    c->S = FN(dup)(sens, (size_t)2 * N * nc);
    c->tmp_in = FN(dup)(NULL, 2 * (size_t)N);
    c->tmp_out = FN(dup)(NULL, 2 * (size_t)N);
    FN(orc_op)* op = (FN(orc_op)*)calloc(1, sizeof(*op));
    op->kind = 3;
    op->n_in = N;
    op->n_out = M * nc;
    op->ctx = c;
    return op;
}

void* FN(orc_dense_create)(const R* A, int rows, int cols)
{
    FN(dense_ctx)* c = (FN(dense_ctx)*)calloc(1, sizeof(*c));
    c->rows = rows;
    c->cols = cols;
    c->A = FN(dup)(A, (size_t)2 * rows * cols);
    FN(orc_op)* op = (FN(orc_op)*)calloc(1, sizeof(*op));
    op->kind = 4;
    op->n_in = cols;
    op->n_out = rows;
    op->ctx = c;
    return op;
}

/* pcSenseTimeSeg.cpp:40-124.  kx, ky, kz, t: Nd*Ns, shot-major (sample d of
 * shot s at s*Nd + d); sens: Ni*Nc complex coil-major; fm: Ni; phase: Ni*Ns
 * real, shot-major.  Gnufft per shot with gridos 2 (:98), TimeSegmentation
 * per shot with shots = 1 (:99-100).  The shot maps (:104-115) are formed in
 * double and rounded once. */
void* FN(orc_pcsense_create)(int Nd, int Ns, const R* kx, const R* ky, const R* kz, int Nx, int Ny,
                             int Nz, int Nc, const R* t, int L, int interptype, const R* sens,
                             const R* fm, const R* phase, R kernelWidth)
{
    FN(pcsense_ctx)* c = (FN(pcsense_ctx)*)calloc(1, sizeof(*c));
    const int Ni = Nx * Ny * Nz;
    c->Ns = Ns;
    c->Nc = Nc;
    c->Nd = Nd;
    c->Ni = Ni;
    c->G = (FN(orc_op)**)calloc((size_t)Ns, sizeof(void*));
    c->A = (FN(orc_op)**)calloc((size_t)Ns, sizeof(void*));
    c->tmp_in = FN(dup)(NULL, 2 * (size_t)Ni);
    c->tmp_out = FN(dup)(NULL, 2 * (size_t)Ni);
    for (int s = 0; s < Ns; ++s) {
        const size_t o = (size_t)s * Nd;
        c->G[s] = (FN(orc_op)*)FN(orc_gnufft_create)(Nd, (R)2.0, Nx, Ny, Nz, kx + o, ky + o,
                                                      kz + o, kernelWidth);
        c->A[s] = (FN(orc_op)*)FN(orc_ts_create)(c->G[s], fm, t + o, Nd, Ni, L, interptype, 1);
    }
    c->SS = FN(dup)(NULL, (size_t)2 * Ni * Nc * Ns);
    c->CS = FN(dup)(NULL, (size_t)2 * Ni * Nc * Ns);
    for (int cc = 0; cc < Nc; ++cc)
        for (int s = 0; s < Ns; ++s) {
            const size_t col = (size_t)s + (size_t)cc * Ns;
            for (int n = 0; n < Ni; ++n) {
                const double sr = sens[2 * ((size_t)cc * Ni + n)],
                             si = sens[2 * ((size_t)cc * Ni + n) + 1];
                const double ph = phase[(size_t)s * Ni + n];
                const double er = cos(ph), ei = -sin(ph); /* exp(-i P) */
                c->SS[2 * (col * Ni + n)] = (R)(sr * er - si * ei);
                c->SS[2 * (col * Ni + n) + 1] = (R)(sr * ei + si * er);
                /* conj(S) conj(exp(-i P)) */
                c->CS[2 * (col * Ni + n)] = (R)(sr * er - si * ei);
                c->CS[2 * (col * Ni + n) + 1] = (R)(-(sr * ei + si * er));
            }
        }
    FN(orc_op)* op = (FN(orc_op)*)calloc(1, sizeof(*op));
    op->kind = 5;
    op->n_in = Ni;
    op->n_out = Nd * Ns * Nc;
    op->ctx = c;
    return op;
}

/* MPI/mpipcSENSETimeSeg.cpp:39-150 for rank `rank` of `nranks`: the tasks
 * t = 0..Ns*Nc-1 with shot t / Nc and coil t % Nc (:106-113) are dealt round
 * robin, t to rank t mod nranks (:115-124); the rank builds, per task, a
 * Gnufft(Nd, 2.0, ...) over the task's shot and a TimeSegmentation over the
 * shot's times with shots = Ns (:143-148).  The forward returns the rank's
 * Nd x ntask columns in task order (tempOutData, :163-171, before the
 * gather); the adjoint takes those columns (column jj = the full data's
 * column shot + coil*Ns, :221-223) and returns the rank's partial image
 * (:218-228, before the reduce). */
void* FN(orc_pcsense_rank_create)(int Nd, int Ns, const R* kx, const R* ky, const R* kz, int Nx,
                                  int Ny, int Nz, int Nc, const R* t, int L, int interptype,
                                  const R* sens, const R* fm, const R* phase, R kernelWidth,
                                  int rank, int nranks)
{
    FN(mpipc_ctx)* c = (FN(mpipc_ctx)*)calloc(1, sizeof(*c));
    const int Ni = Nx * Ny * Nz;
    c->Ni = Ni;
    c->Nd = Nd;
    int n = 0;
    for (int task = rank; task < Ns * Nc; task += nranks) ++n;
    c->ntask = n;
    c->shot = (int*)calloc((size_t)n + 1, sizeof(int));
    c->coil = (int*)calloc((size_t)n + 1, sizeof(int));
    c->G = (FN(orc_op)**)calloc((size_t)n + 1, sizeof(void*));
    c->A = (FN(orc_op)**)calloc((size_t)n + 1, sizeof(void*));
    int jj = 0;
    for (int task = rank; task < Ns * Nc; task += nranks, ++jj) {
        c->shot[jj] = task / Nc;
        c->coil[jj] = task % Nc;
        const size_t o = (size_t)c->shot[jj] * Nd;
        c->G[jj] = (FN(orc_op)*)FN(orc_gnufft_create)(Nd, (R)2.0, Nx, Ny, Nz, kx + o, ky + o,
                                                       kz + o, kernelWidth);
        c->A[jj] = (FN(orc_op)*)FN(orc_ts_create)(c->G[jj], fm, t + o, Nd, Ni, L, interptype, Ns);
    }
    c->S = FN(dup)(sens, (size_t)2 * Ni * Nc);
    c->P = FN(dup)(phase, (size_t)Ni * Ns);
    c->tmp_in = FN(dup)(NULL, 2 * (size_t)Ni);
    c->tmp_out = FN(dup)(NULL, 2 * (size_t)Ni);
    FN(orc_op)* op = (FN(orc_op)*)calloc(1, sizeof(*op));
    op->kind = 7;
    op->n_in = Ni;
    op->n_out = Nd * n;
    op->ctx = c;
    return op;
}

/* the task's map S_c exp(-i P_s) at voxel v, in R */
static void FN(mpipc_map)(const FN(mpipc_ctx) * c, int jj, int v, R* mr, R* mi)
{
    const R* sp = c->S + 2 * ((size_t)c->coil[jj] * c->Ni + v);
    const R ph = c->P[(size_t)c->shot[jj] * c->Ni + v];
    const R er = (R)cos((double)ph), ei = (R)-sin((double)ph);
    *mr = sp[0] * er - sp[1] * ei;
    *mi = sp[0] * ei + sp[1] * er;
}

int FN(orc_pcsense_rank_tasks)(void* op_, int* shot, int* coil)
{
    FN(orc_op)* op = (FN(orc_op)*)op_;
    if (op->kind != 7) return -1;
    FN(mpipc_ctx)* c = (FN(mpipc_ctx)*)op->ctx;
    for (int jj = 0; jj < c->ntask; ++jj) {
        if (shot) shot[jj] = c->shot[jj];
        if (coil) coil[jj] = c->coil[jj];
    }
    return c->ntask;
}

/* pcSENSE.cpp:39-102: the same shot/coil structure over an exact DFT per
 * shot (Gdft with the shot's samples and times, the field map, image
 * coordinates (x - N/2)/N formed in double, y fastest, pcSENSE.cpp:62-80).
 * The reference's maps: forward S_c exp(+i P_s) (expiPMap = conj(exp(-i P)),
 * :97, used at :117), adjoint conj(S_c) exp(+i P_s) (:98, :142), so the two
 * operators are not each other's adjoint; restated as written. */
void* FN(orc_pcsense_dft_create)(int Nd, int Ns, const R* kx, const R* ky, const R* kz, int Nx,
                                 int Ny, int Nz, int Nc, const R* t, const R* sens, const R* fm,
                                 const R* phase)
{
    FN(pcsense_ctx)* c = (FN(pcsense_ctx)*)calloc(1, sizeof(*c));
    const int Ni = Nx * Ny * Nz;
    c->Ns = Ns;
    c->Nc = Nc;
    c->Nd = Nd;
    c->Ni = Ni;
    c->G = (FN(orc_op)**)calloc((size_t)Ns, sizeof(void*));
    c->A = (FN(orc_op)**)calloc((size_t)Ns, sizeof(void*));
    c->tmp_in = FN(dup)(NULL, 2 * (size_t)Ni);
    c->tmp_out = FN(dup)(NULL, 2 * (size_t)Ni);
    R* ix = FN(dup)(NULL, (size_t)Ni);
    R* iy = FN(dup)(NULL, (size_t)Ni);
    R* iz = FN(dup)(NULL, (size_t)Ni);
    for (int z = 0; z < Nz; ++z)
        for (int x = 0; x < Nx; ++x)
            for (int y = 0; y < Ny; ++y) {
                const size_t n = (size_t)y + (size_t)x * Ny + (size_t)z * Nx * Ny;
                ix[n] = (R)(((double)(R)x - (double)(R)Nx / 2.0) / (double)(R)Nx);
                iy[n] = (R)(((double)(R)y - (double)(R)Ny / 2.0) / (double)(R)Ny);
                iz[n] = (R)(((double)(R)z - (double)(R)Nz / 2.0) / (double)(R)Nz);
            }
    for (int s = 0; s < Ns; ++s) {
        const size_t o = (size_t)s * Nd;
        c->A[s] = (FN(orc_op)*)FN(orc_gdft_create)(Nd, Ni, kx + o, ky + o, kz + o, ix, iy, iz,
                                                    fm, t + o);
    }
    free(ix);
    free(iy);
    free(iz);
    c->SS = FN(dup)(NULL, (size_t)2 * Ni * Nc * Ns);
    c->CS = FN(dup)(NULL, (size_t)2 * Ni * Nc * Ns);
    for (int cc = 0; cc < Nc; ++cc)
        for (int s = 0; s < Ns; ++s) {
            const size_t col = (size_t)s + (size_t)cc * Ns;
            for (int n = 0; n < Ni; ++n) {
                const double sr = sens[2 * ((size_t)cc * Ni + n)],
                             si = sens[2 * ((size_t)cc * Ni + n) + 1];
                const double ph = phase[(size_t)s * Ni + n];
                const double er = cos(ph), ei = sin(ph); /* exp(+i P) */
                c->SS[2 * (col * Ni + n)] = (R)(sr * er - si * ei);
                c->SS[2 * (col * Ni + n) + 1] = (R)(sr * ei + si * er);
                /* conj(S) exp(+i P) */
                c->CS[2 * (col * Ni + n)] = (R)(sr * er + si * ei);
                c->CS[2 * (col * Ni + n) + 1] = (R)(sr * ei - si * er);
            }
        }
    FN(orc_op)* op = (FN(orc_op)*)calloc(1, sizeof(*op));
    op->kind = 5;
    op->n_in = Ni;
    op->n_out = Nd * Ns * Nc;
    op->ctx = c;
    return op;
}

static inline void FN(cmul)(R ar, R ai, R br, R bi, R* cr, R* ci)
{
    *cr = ar * br - ai * bi;
    *ci = ar * bi + ai * br;
}

void FN(orc_op_forward)(void* opv, const R* in, R* out)
{
    FN(orc_op)* op = (FN(orc_op)*)opv;
    switch (op->kind) {
    case 0:
        FN(nufft_fwd)((FN(nufft_ctx)*)op->ctx, in, out);
        break;
    case 1: {
        FN(dft_ctx)* c = (FN(dft_ctx)*)op->ctx;
        if (c->gx)
            FN(ft_grads)(c, in, out);
        else
            FN(orc_ft)(out, in, c->kx, c->ky, c->kz, c->ix, c->iy, c->iz, c->fm, c->t, c->M, c->N);
        break;
    }
    case 2: {
        /* TimeSegmentation.cpp:252-296 */
        FN(ts_ctx)* c = (FN(ts_ctx)*)op->ctx;
        memset(out, 0, sizeof(R) * 2 * c->M);
        for (int l = 0; l < c->L; ++l) {
            const R* wo = c->Wo + 2 * (size_t)l * c->N;
            for (int n = 0; n < c->N; ++n)
                FN(cmul)(wo[2 * n], wo[2 * n + 1], in[2 * n], in[2 * n + 1], &c->tmp_in[2 * n],
                         &c->tmp_in[2 * n + 1]);
            FN(orc_op_forward)(c->G, c->tmp_in, c->tmp_out);
            if (c->AAc) {
                const R* ac = c->AAc + 2 * (size_t)l * c->M;
                for (int j = 0; j < c->M; ++j) {
                    R cr, ci;
                    FN(cmul)(c->tmp_out[2 * j], c->tmp_out[2 * j + 1], ac[2 * j], ac[2 * j + 1],
                             &cr, &ci);
                    out[2 * j] += cr;
                    out[2 * j + 1] += ci;
                }
                continue;
            }
            const R* aa = c->AA + (size_t)l * c->M;
            for (int j = 0; j < c->M; ++j) {
                out[2 * j] += aa[j] * c->tmp_out[2 * j];
                out[2 * j + 1] += aa[j] * c->tmp_out[2 * j + 1];
            }
        }
        break;
    }
    case 3: {
        /* SENSE.cpp:51-71 */
        FN(sense_ctx)* c = (FN(sense_ctx)*)op->ctx;
        for (int cc = 0; cc < c->nc; ++cc) {
            const R* s = c->S + 2 * (size_t)cc * c->N;
            for (int n = 0; n < c->N; ++n)
                FN(cmul)(s[2 * n], s[2 * n + 1], in[2 * n], in[2 * n + 1], &c->tmp_in[2 * n],
                         &c->tmp_in[2 * n + 1]);
            FN(orc_op_forward)(c->G, c->tmp_in, out + 2 * (size_t)cc * c->M);
        }
        break;
    }
    case 7: {
        /* MPI/mpipcSENSETimeSeg.cpp:163-171 */
        FN(mpipc_ctx)* c = (FN(mpipc_ctx)*)op->ctx;
        for (int jj = 0; jj < c->ntask; ++jj) {
            for (int n = 0; n < c->Ni; ++n) {
                R mr, mi;
                FN(mpipc_map)(c, jj, n, &mr, &mi);
                FN(cmul)(in[2 * n], in[2 * n + 1], mr, mi, &c->tmp_in[2 * n], &c->tmp_in[2 * n + 1]);
            }
            FN(orc_op_forward)(c->A[jj], c->tmp_in, out + 2 * (size_t)jj * c->Nd);
        }
        break;
    }
    case 5: {
        /* pcSenseTimeSeg.cpp:129-148: column s + c*Ns of an Nd x (Ns*Nc) array */
        FN(pcsense_ctx)* c = (FN(pcsense_ctx)*)op->ctx;
        for (int cc = 0; cc < c->Nc; ++cc)
            for (int s = 0; s < c->Ns; ++s) {
                const size_t col = (size_t)s + (size_t)cc * c->Ns;
                const R* m = c->SS + 2 * col * c->Ni;
                for (int n = 0; n < c->Ni; ++n)
                    FN(cmul)(in[2 * n], in[2 * n + 1], m[2 * n], m[2 * n + 1], &c->tmp_in[2 * n],
                             &c->tmp_in[2 * n + 1]);
                FN(orc_op_forward)(c->A[s], c->tmp_in, out + 2 * col * c->Nd);
            }
        break;
    }
    case 4: {
        FN(dense_ctx)* c = (FN(dense_ctx)*)op->ctx;
        for (int r = 0; r < c->rows; ++r) {
            R sr = 0, si = 0;
            for (int q = 0; q < c->cols; ++q) {
                R cr, ci;
                const R* a = c->A + 2 * ((size_t)q * c->rows + r);
                FN(cmul)(a[0], a[1], in[2 * q], in[2 * q + 1], &cr, &ci);
                sr += cr;
                si += ci;
            }
            out[2 * r] = sr;
            out[2 * r + 1] = si;
        }
        break;
    }
    }
}

void FN(orc_op_adjoint)(void* opv, const R* in, R* out)
{
    FN(orc_op)* op = (FN(orc_op)*)opv;
    switch (op->kind) {
    case 0:
        FN(nufft_adj)((FN(nufft_ctx)*)op->ctx, in, out);
        break;
    case 1: {
        FN(dft_ctx)* c = (FN(dft_ctx)*)op->ctx;
        if (c->gx)
            FN(ift_grads)(c, in, out);
        else
            FN(orc_ift)(out, in, c->kx, c->ky, c->kz, c->ix, c->iy, c->iz, c->fm, c->t, c->M, c->N);
        break;
    }
    case 2: {
        /* TimeSegmentation.cpp:297-324 */
        FN(ts_ctx)* c = (FN(ts_ctx)*)op->ctx;
        memset(out, 0, sizeof(R) * 2 * c->N);
        for (int l = 0; l < c->L; ++l) {
            const R* aa = c->AA + (size_t)l * c->M;
            const R* ac = c->AAc ? c->AAc + 2 * (size_t)l * c->M : NULL;
            for (int j = 0; j < c->M; ++j) {
                if (ac) {
                    /* conj(AA) * d */
                    FN(cmul)(ac[2 * j], -ac[2 * j + 1], in[2 * j], in[2 * j + 1],
                             &c->tmp_in[2 * j], &c->tmp_in[2 * j + 1]);
                } else {
                    /* conj(AA) * d; AA is real-valued */
                    c->tmp_in[2 * j] = aa[j] * in[2 * j];
                    c->tmp_in[2 * j + 1] = aa[j] * in[2 * j + 1];
                }
            }
            FN(orc_op_adjoint)(c->G, c->tmp_in, c->tmp_out);
            const R* woh = c->WoH + 2 * (size_t)l * c->N;
            for (int n = 0; n < c->N; ++n) {
                R cr, ci;
                FN(cmul)(c->tmp_out[2 * n], c->tmp_out[2 * n + 1], woh[2 * n], woh[2 * n + 1], &cr,
                         &ci);
                out[2 * n] += cr;
                out[2 * n + 1] += ci;
            }
        }
        break;
    }
    case 3: {
        /* SENSE.cpp:75-97 */
        FN(sense_ctx)* c = (FN(sense_ctx)*)op->ctx;
        memset(out, 0, sizeof(R) * 2 * c->N);
        for (int cc = 0; cc < c->nc; ++cc) {
            FN(orc_op_adjoint)(c->G, in + 2 * (size_t)cc * c->M, c->tmp_out);
            const R* s = c->S + 2 * (size_t)cc * c->N;
            for (int n = 0; n < c->N; ++n) {
                R cr, ci;
                FN(cmul)(c->tmp_out[2 * n], c->tmp_out[2 * n + 1], s[2 * n], -s[2 * n + 1], &cr,
                         &ci);
                out[2 * n] += cr;
                out[2 * n + 1] += ci;
            }
        }
        break;
    }
    case 7: {
        /* MPI/mpipcSENSETimeSeg.cpp:218-228: sum over the tasks, in task order,
           of conj(S_c exp(-i P_s)) (x) A_jj' d[:, jj] */
        FN(mpipc_ctx)* c = (FN(mpipc_ctx)*)op->ctx;
        memset(out, 0, sizeof(R) * 2 * c->Ni);
        for (int jj = 0; jj < c->ntask; ++jj) {
            FN(orc_op_adjoint)(c->A[jj], in + 2 * (size_t)jj * c->Nd, c->tmp_out);
            for (int n = 0; n < c->Ni; ++n) {
                R mr, mi, cr, ci;
                FN(mpipc_map)(c, jj, n, &mr, &mi);
                FN(cmul)(mr, -mi, c->tmp_out[2 * n], c->tmp_out[2 * n + 1], &cr, &ci);
                out[2 * n] += cr;
                out[2 * n + 1] += ci;
            }
        }
        break;
    }
    case 5: {
        /* pcSenseTimeSeg.cpp:152-171: sum over coils (outer) and shots (inner) of
           CS[c*Ns + s] (x) A_s' d[:, s + c*Ns] */
        FN(pcsense_ctx)* c = (FN(pcsense_ctx)*)op->ctx;
        memset(out, 0, sizeof(R) * 2 * c->Ni);
        for (int cc = 0; cc < c->Nc; ++cc)
            for (int s = 0; s < c->Ns; ++s) {
                const size_t col = (size_t)s + (size_t)cc * c->Ns;
                FN(orc_op_adjoint)(c->A[s], in + 2 * col * c->Nd, c->tmp_out);
                const R* m = c->CS + 2 * col * c->Ni;
                for (int n = 0; n < c->Ni; ++n) {
                    R cr, ci;
                    FN(cmul)(m[2 * n], m[2 * n + 1], c->tmp_out[2 * n], c->tmp_out[2 * n + 1], &cr,
                             &ci);
                    out[2 * n] += cr;
                    out[2 * n + 1] += ci;
                }
            }
        break;
    }
    case 4: {
        FN(dense_ctx)* c = (FN(dense_ctx)*)op->ctx;
        for (int q = 0; q < c->cols; ++q) {
            R sr = 0, si = 0;
            for (int r = 0; r < c->rows; ++r) {
                R cr, ci;
                const R* a = c->A + 2 * ((size_t)q * c->rows + r);
                FN(cmul)(a[0], -a[1], in[2 * r], in[2 * r + 1], &cr, &ci);
                sr += cr;
                si += ci;
            }
            out[2 * q] = sr;
            out[2 * q + 1] = si;
        }
        break;
    }
    }
}

int FN(orc_op_n_in)(void* op) { return ((FN(orc_op)*)op)->n_in; }
int FN(orc_op_n_out)(void* op) { return ((FN(orc_op)*)op)->n_out; }

void FN(orc_op_destroy)(void* opv)
{
    FN(orc_op)* op = (FN(orc_op)*)opv;
    if (!op) return;
    switch (op->kind) {
    case 0: {
        FN(nufft_ctx)* c = (FN(nufft_ctx)*)op->ctx;
        free(c->lut); free(c->kx); free(c->ky); free(c->kz);
        free(c->img); free(c->img_d); free(c->grid); free(c->grid_d);
        break;
    }
    case 1: {
        FN(dft_ctx)* c = (FN(dft_ctx)*)op->ctx;
        free(c->kx); free(c->ky); free(c->kz); free(c->t);
        free(c->ix); free(c->iy); free(c->iz); free(c->fm);
        free(c->gx); free(c->gy); free(c->gz);
        break;
    }
    case 2: {
        FN(ts_ctx)* c = (FN(ts_ctx)*)op->ctx;
        free(c->AA); free(c->AAc); free(c->Wo); free(c->WoH);
        free(c->tmp_in); free(c->tmp_out); free(c->acc);
        break;
    }
    case 3: {
        FN(sense_ctx)* c = (FN(sense_ctx)*)op->ctx;
        free(c->S); free(c->tmp_in); free(c->tmp_out);
        break;
    }
    case 4: {
        FN(dense_ctx)* c = (FN(dense_ctx)*)op->ctx;
        free(c->A);
        break;
    }
    case 7: {
        FN(mpipc_ctx)* c = (FN(mpipc_ctx)*)op->ctx;
        for (int jj = 0; jj < c->ntask; ++jj) {
            FN(orc_op_destroy)(c->A[jj]);
            FN(orc_op_destroy)(c->G[jj]);
        }
        free(c->A); free(c->G); free(c->shot); free(c->coil); free(c->S); free(c->P);
        free(c->tmp_in); free(c->tmp_out);
        break;
    }
    case 5: {
        FN(pcsense_ctx)* c = (FN(pcsense_ctx)*)op->ctx;
        for (int s = 0; s < c->Ns; ++s) {
            FN(orc_op_destroy)(c->A[s]);
            FN(orc_op_destroy)(c->G[s]);
        }
        free(c->A); free(c->G); free(c->SS); free(c->CS); free(c->tmp_in); free(c->tmp_out);
        break;
    }
    }
    free(op->ctx);
    free(op);
}

/* --------------------------------------------------------- penalties -- */

typedef struct {
    int kind; /* 0 quadratic, 1 TV */
    int Nx, Ny, Nz, dims2pen;
    R beta, delta;
} FN(pen_ctx);

void* FN(orc_penalty_create)(int kind, int Nx, int Ny, int Nz, R beta, R delta, int dims2pen)
{
    FN(pen_ctx)* p = (FN(pen_ctx)*)calloc(1, sizeof(*p));
    p->kind = kind;
    p->Nx = Nx;
    p->Ny = Ny;
    p->Nz = Nz;
    p->beta = beta;
    p->delta = delta;
    p->dims2pen = dims2pen;
    return p;
}

void FN(orc_penalty_destroy)(void* p) { free(p); }

static long FN(pen_offset)(FN(pen_ctx) * p, int dim)
{
    /* Robject.cpp:55-76 */
    long ll = 0, jj = 0, kk = 0;
    if (dim == 0) ll = 1;
    else if (dim == 1) jj = 1;
    else kk = 1;
    return ll + jj * p->Ny + kk * (long)p->Nx * p->Ny;
}

/* Robject.cpp:49-82 */
void FN(orc_penalty_cd)(void* pv, const R* d, int dim, R* out)
{
    FN(pen_ctx)* p = (FN(pen_ctx)*)pv;
    long n = (long)p->Nx * p->Ny * p->Nz;
    long off = FN(pen_offset)(p, dim);
    memset(out, 0, sizeof(R) * 2 * n);
    for (long ii = off; ii < n; ii++) {
        out[2 * ii] = d[2 * ii] - d[2 * (ii - off)];
        out[2 * ii + 1] = d[2 * ii + 1] - d[2 * (ii - off) + 1];
    }
}

/* Robject.cpp:84-120 */
void FN(orc_penalty_ctd)(void* pv, const R* d, int dim, R* out)
{
    FN(pen_ctx)* p = (FN(pen_ctx)*)pv;
    long n = (long)p->Nx * p->Ny * p->Nz;
    long off = FN(pen_offset)(p, dim);
    memset(out, 0, sizeof(R) * 2 * n);
    out[2 * (off - 1)] = -d[2 * off];
    out[2 * (off - 1) + 1] = -d[2 * off + 1];
    out[2 * (n - 1 - off)] = d[2 * (n - 1 - off)];
    out[2 * (n - 1 - off) + 1] = d[2 * (n - 1 - off) + 1];
    for (long ii = off; ii < n - 1; ii++) {
        out[2 * (ii - off)] = d[2 * (ii - off)] - d[2 * ii];
        out[2 * (ii - off) + 1] = d[2 * (ii - off) + 1] - d[2 * ii + 1];
    }
}

/* QuadPenalty.cpp:45-64, TVPenalty.cpp:43-68: wpot, dpot applied in place */
static void FN(pen_wpot)(FN(pen_ctx) * p, R* d, long n)
{
    for (long i = 0; i < n; ++i) {
        R w = 1;
        if (p->kind == 1) {
            R ar = d[2 * i] / p->delta, ai = d[2 * i + 1] / p->delta;
            R t = (R)sqrt(ar * ar + ai * ai);
            w = (R)1.0 / (R)sqrt((R)1.0 + t * t);
        }
        d[2 * i] = w;
        d[2 * i + 1] = 0;
    }
}

static void FN(pen_dpot)(FN(pen_ctx) * p, R* d, long n)
{
    if (p->kind != 1) return;
    for (long i = 0; i < n; ++i) {
        R ar = d[2 * i] / p->delta, ai = d[2 * i + 1] / p->delta;
        R t = (R)sqrt(ar * ar + ai * ai);
        R s = (R)sqrt((R)1.0 + t * t);
        d[2 * i] = d[2 * i] / s;
        d[2 * i + 1] = d[2 * i + 1] / s;
    }
}

static int FN(pen_nd)(FN(pen_ctx) * p)
{
    /* Robject.cpp:128-133 */
    return ((p->Nz == 1) || (p->dims2pen == 2)) ? 2 : 3;
}

/* Robject.cpp:143-165 */
void FN(orc_penalty_gradient)(void* pv, const R* x, R* g)
{
    FN(pen_ctx)* p = (FN(pen_ctx)*)pv;
    long n = (long)p->Nx * p->Ny * p->Nz;
    R* d = FN(dup)(NULL, 2 * n);
    R* e = FN(dup)(NULL, 2 * n);
    memset(g, 0, sizeof(R) * 2 * n);
    for (int dim = 0; dim < FN(pen_nd)(p); ++dim) {
        FN(orc_penalty_cd)(p, x, dim, d);
        FN(pen_dpot)(p, d, n);
        FN(orc_penalty_ctd)(p, d, dim, e);
        for (long i = 0; i < 2 * n; ++i) g[i] += e[i];
    }
    for (long i = 0; i < 2 * n; ++i) g[i] = p->beta * g[i];
    free(d);
    free(e);
}

/* Robject.cpp:167-193 (returns a complex number in out2[0..1]) */
void FN(orc_penalty_denom)(void* pv, const R* ddir, const R* x, R* out2)
{
    FN(pen_ctx)* p = (FN(pen_ctx)*)pv;
    long n = (long)p->Nx * p->Ny * p->Nz;
    R* cdir = FN(dup)(NULL, 2 * n);
    R* cx = FN(dup)(NULL, 2 * n);
    R pr = 0, pi = 0;
    for (int dim = 0; dim < FN(pen_nd)(p); ++dim) {
        FN(orc_penalty_cd)(p, ddir, dim, cdir);
        FN(orc_penalty_cd)(p, x, dim, cx);
        FN(pen_wpot)(p, cx, n);
        R tr = 0, ti = 0;
        for (long i = 0; i < n; ++i) {
            R wr, wi; /* Cx %= Cdir */
            FN(cmul)(cx[2 * i], cx[2 * i + 1], cdir[2 * i], cdir[2 * i + 1], &wr, &wi);
            R cr, ci; /* cdot(Cdir, Cx) = sum conj(Cdir) * Cx */
            FN(cmul)(cdir[2 * i], -cdir[2 * i + 1], wr, wi, &cr, &ci);
            tr += cr;
            ti += ci;
        }
        pr += tr;
        pi += ti;
    }
    out2[0] = pr * p->beta - pi * 0;
    out2[1] = pr * 0 + pi * p->beta;
    free(cdir);
    free(cx);
}

/* Robject.cpp:122-142 */
R FN(orc_penalty_value)(void* pv, const R* x)
{
    FN(pen_ctx)* p = (FN(pen_ctx)*)pv;
    long n = (long)p->Nx * p->Ny * p->Nz;
    R* d = FN(dup)(NULL, 2 * n);
    R penal = 0;
    for (int dim = 0; dim < FN(pen_nd)(p); ++dim) {
        FN(orc_penalty_cd)(p, x, dim, d);
        R s = 0;
        for (long i = 0; i < n; ++i) {
            R a2 = d[2 * i] * d[2 * i] + d[2 * i + 1] * d[2 * i + 1];
            if (p->kind == 1)
                s += p->delta * p->delta *
                     ((R)sqrt((R)1.0 + a2 / (p->delta * p->delta)) - (R)1.0);
            else s += a2 / (R)2.0;
        }
        penal = penal + (R)fabs(s);
    }
    free(d);
    return p->beta * penal;
}

/* --------------------------------------------------------------- PCG -- */

static void FN(cdot)(const R* a, const R* b, long n, R* re, R* im)
{
    R sr = 0, si = 0;
    for (long i = 0; i < n; ++i) {
        sr += a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
        si += a[2 * i] * b[2 * i + 1] - a[2 * i + 1] * b[2 * i];
    }
    *re = sr;
    *im = si;
}

/* solve_pwls_pcg.hpp:48-167.  Returns the number of completed updates.
 * Scalars that the reference keeps complex are real-valued on this path
 * (every one is formed from real(...) or from real-valued penalties); the
 * imaginary part of the penalty denominator is carried through. */
int FN(orc_pcg)(void* opv, void* Rv, const R* x0, const R* W, const R* yi, int niter,
                R* x_out, R* resid)
{
    FN(orc_op)* A = (FN(orc_op)*)opv;
    long N = A->n_in, M = A->n_out;
    R* x = FN(dup)(x0, 2 * N);
    R* Ax = FN(dup)(NULL, 2 * M);
    R* r = FN(dup)(NULL, 2 * M);
    R* ngrad = FN(dup)(NULL, 2 * N);
    R* pg = FN(dup)(NULL, 2 * N);
    R* ddir = FN(dup)(NULL, 2 * N);
    R* Adir = FN(dup)(NULL, 2 * M);
    R* xs = FN(dup)(NULL, 2 * N);
    int done = 0;
    FN(orc_op_forward)(A, x, Ax);
    R oldinprod = 0, gamma = 0;
    /* yi' W yi */
    R ywy = 0;
    for (long i = 0; i < M; ++i) {
        R w = W ? W[i] : (R)1;
        ywy += w * (yi[2 * i] * yi[2 * i] + yi[2 * i + 1] * yi[2 * i + 1]);
    }
    for (int ii = 0; ii < niter; ii++) {
        for (long i = 0; i < M; ++i) {
            R w = W ? W[i] : (R)1;
            r[2 * i] = w * (yi[2 * i] - Ax[2 * i]);
            r[2 * i + 1] = w * (yi[2 * i + 1] - Ax[2 * i + 1]);
        }
        FN(orc_op_adjoint)(A, r, ngrad);
        R nrm = 0;
        for (long i = 0; i < 2 * N; ++i) nrm += ngrad[i] * ngrad[i];
        nrm = (R)sqrt(nrm);
        if (nrm / ywy < 1e-10) break; /* :88-91 */
        if (Rv) {
            FN(orc_penalty_gradient)(Rv, x, pg);
            for (long i = 0; i < 2 * N; ++i) ngrad[i] -= pg[i];
        }
        R newinprod, dummy;
        FN(cdot)(ngrad, ngrad, N, &newinprod, &dummy);
        if (ii == 0) {
            memcpy(ddir, ngrad, sizeof(R) * 2 * N);
        } else {
            if (fabs(oldinprod) < 1e-10) gamma = 0;
            else gamma = newinprod / oldinprod;
            for (long i = 0; i < 2 * N; ++i) ddir[i] = ngrad[i] + gamma * ddir[i];
        }
        oldinprod = newinprod;
        R dg;
        FN(cdot)(ddir, ngrad, N, &dg, &dummy);
        if (dg < 0) break; /* :113-116 */
        FN(orc_op_forward)(A, ddir, Adir);
        R dAWAd = 0, dAWr = 0;
        for (long i = 0; i < M; ++i) {
            R w = W ? W[i] : (R)1;
            dAWAd += Adir[2 * i] * w * Adir[2 * i] + Adir[2 * i + 1] * w * Adir[2 * i + 1];
            R rr = w * (yi[2 * i] - Ax[2 * i]), ri = w * (yi[2 * i + 1] - Ax[2 * i + 1]);
            dAWr += Adir[2 * i] * rr + Adir[2 * i + 1] * ri;
        }
        R step = 0;
        int bad = 0;
        for (int j = 0; j < 3; j++) {
            for (long i = 0; i < 2 * N; ++i) xs[i] = x[i] + step * ddir[i];
            R pden[2] = {0, 0};
            if (Rv) FN(orc_penalty_denom)(Rv, ddir, xs, pden);
            R dr = dAWAd + pden[0], di = pden[1];
            /* std::abs of the complex denom: hypot, no overflow of dr^2 in float */
            R ad = (R)hypot(dr, di);
            if (ad < 1e-20 || ad > 1e25) { bad = 1; break; } /* :137-145 */
            R pdot = 0;
            if (Rv) {
                FN(orc_penalty_gradient)(Rv, xs, pg);
                FN(cdot)(ddir, pg, N, &pdot, &dummy);
            }
            /* step -= (-dAWr + step*dAWAd + pdot) / denom (complex denom),
               divided with scaling as complex division does */
            R num = -dAWr + step * dAWAd + pdot;
            R sdr = dr / ad, sdi = di / ad;
            step -= (num / ad) * sdr / (sdr * sdr + sdi * sdi);
        }
        if (bad) break;
        for (long i = 0; i < 2 * M; ++i) Ax[i] += step * Adir[i];
        for (long i = 0; i < 2 * N; ++i) x[i] += step * ddir[i];
        done++;
        if (resid) {
            R s = 0;
            for (long i = 0; i < M; ++i) {
                R a = yi[2 * i] - Ax[2 * i], b = yi[2 * i + 1] - Ax[2 * i + 1];
                s += a * a + b * b;
            }
            resid[ii] = (R)sqrt(s);
        }
    }
    memcpy(x_out, x, sizeof(R) * 2 * N);
    free(x); free(Ax); free(r); free(ngrad); free(pg); free(ddir); free(Adir); free(xs);
    return done;
}

#undef FN
