/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Per-precision body of the device-order CPU replay, instantiated by
 * dev_exact.c with REAL = float (RFMA = fmaf) and REAL = double (RFMA = fma).
 * Storage: links[((mu*2 + parity)*NC + c)*Vh + slot].
 */
#define YM_CAT2(a, b) a##_##b
#define YM_CAT(a, b) YM_CAT2(a, b)
#define F(name) YM_CAT(name, RSUF)

static inline REAL *F(lk)(REAL *L, const geom *g, int mu, int p, int nc, int c, int64_t s) {
    return L + (((int64_t)(mu * 2 + p) * nc + c) * g->vh + s);
}
static void F(load_link)(const REAL *L, const geom *g, const int64_t x[4], int mu, int nc, REAL *out) {
    int p = parity_of(x);
    int64_t s = slot_of(g, x);
    for (int c = 0; c < nc; ++c) out[c] = L[((int64_t)(mu * 2 + p) * nc + c) * g->vh + s];
}
static void F(store_link)(REAL *L, const geom *g, const int64_t x[4], int mu, int nc, const REAL *v) {
    int p = parity_of(x);
    int64_t s = slot_of(g, x);
    for (int c = 0; c < nc; ++c) L[((int64_t)(mu * 2 + p) * nc + c) * g->vh + s] = v[c];
}

static int F(to_device)(const int64_t n[4], const double *aos, void *dev, int nc) {
    geom g = make_geom(n);
    REAL *L = (REAL *)dev;
    /* padding slots hold the identity */
    YMD_PARALLEL_FOR
    for (int mu = 0; mu < 4; ++mu) for (int p = 0; p < 2; ++p) for (int c = 0; c < nc; ++c) for (int64_t s = 0; s < g.vh; ++s) {
        REAL v = 0;
        if (nc == 18 && (c == 0 || c == 8 || c == 16)) v = 1;
        if (nc == 4 && c == 0) v = 1;
        *F(lk)(L, &g, mu, p, nc, c, s) = v;
    }
    YMD_PARALLEL_FOR
    for (int64_t x0 = 0; x0 < n[0]; ++x0)
    for (int64_t x1 = 0; x1 < n[1]; ++x1) for (int64_t x2 = 0; x2 < n[2]; ++x2) for (int64_t x3 = 0; x3 < n[3]; ++x3) {
        const int64_t x[4] = {x0, x1, x2, x3};
        int64_t site = ((x[0] * n[1] + x[1]) * n[2] + x[2]) * n[3] + x[3];
        for (int mu = 0; mu < 4; ++mu) {
            REAL v[18];
            for (int c = 0; c < nc; ++c) v[c] = (REAL)aos[(site * 4 + mu) * nc + c];
            F(store_link)(L, &g, x, mu, nc, v);
        }
    }
    return YMO_OK;
}
static int F(from_device)(const int64_t n[4], const void *dev, double *aos, int nc) {
    geom g = make_geom(n);
    const REAL *L = (const REAL *)dev;
    YMD_PARALLEL_FOR
    for (int64_t x0 = 0; x0 < n[0]; ++x0)
    for (int64_t x1 = 0; x1 < n[1]; ++x1) for (int64_t x2 = 0; x2 < n[2]; ++x2) for (int64_t x3 = 0; x3 < n[3]; ++x3) {
        const int64_t x[4] = {x0, x1, x2, x3};
        int64_t site = ((x[0] * n[1] + x[1]) * n[2] + x[2]) * n[3] + x[3];
        for (int mu = 0; mu < 4; ++mu) {
            REAL v[18];
            F(load_link)(L, &g, x, mu, nc, v);
            for (int c = 0; c < nc; ++c) aos[(site * 4 + mu) * nc + c] = (double)v[c];
        }
    }
    return YMO_OK;
}

/* ---------------------------------------------------- SU(3) staple (REAL) */
/* element (i,j) of a 3x3 complex matrix stored as 18 reals */
#define RE(m, i, j) (m)[2 * (3 * (i) + (j))]
#define IM(m, i, j) (m)[2 * (3 * (i) + (j)) + 1]

/* T = X Y^dagger */
static void F(mul_nd)(const REAL *X, const REAL *Y, REAL *T) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
        REAL re = RE(X, i, 0) * RE(Y, j, 0);
        re = RFMA(IM(X, i, 0), IM(Y, j, 0), re);
        re = RFMA(RE(X, i, 1), RE(Y, j, 1), re);
        re = RFMA(IM(X, i, 1), IM(Y, j, 1), re);
        re = RFMA(RE(X, i, 2), RE(Y, j, 2), re);
        re = RFMA(IM(X, i, 2), IM(Y, j, 2), re);
        REAL im = IM(X, i, 0) * RE(Y, j, 0);
        im = RFMA(-RE(X, i, 0), IM(Y, j, 0), im);
        im = RFMA(IM(X, i, 1), RE(Y, j, 1), im);
        im = RFMA(-RE(X, i, 1), IM(Y, j, 1), im);
        im = RFMA(IM(X, i, 2), RE(Y, j, 2), im);
        im = RFMA(-RE(X, i, 2), IM(Y, j, 2), im);
        RE(T, i, j) = re;
        IM(T, i, j) = im;
    }
}
/* T = X^dagger Y^dagger */
static void F(mul_dd)(const REAL *X, const REAL *Y, REAL *T) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
        REAL re = RE(X, 0, i) * RE(Y, j, 0);
        re = RFMA(-IM(X, 0, i), IM(Y, j, 0), re);
        re = RFMA(RE(X, 1, i), RE(Y, j, 1), re);
        re = RFMA(-IM(X, 1, i), IM(Y, j, 1), re);
        re = RFMA(RE(X, 2, i), RE(Y, j, 2), re);
        re = RFMA(-IM(X, 2, i), IM(Y, j, 2), re);
        REAL im = (-RE(X, 0, i)) * IM(Y, j, 0);
        im = RFMA(-IM(X, 0, i), RE(Y, j, 0), im);
        im = RFMA(-RE(X, 1, i), IM(Y, j, 1), im);
        im = RFMA(-IM(X, 1, i), RE(Y, j, 1), im);
        im = RFMA(-RE(X, 2, i), IM(Y, j, 2), im);
        im = RFMA(-IM(X, 2, i), RE(Y, j, 2), im);
        RE(T, i, j) = re;
        IM(T, i, j) = im;
    }
}
/* A += T Z^dagger */
static void F(acc_nd)(REAL *A, const REAL *T, const REAL *Z) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
        REAL re = RE(A, i, j), im = IM(A, i, j);
        for (int k = 0; k < 3; ++k) {
            re = RFMA(RE(T, i, k), RE(Z, j, k), re);
            re = RFMA(IM(T, i, k), IM(Z, j, k), re);
        }
        for (int k = 0; k < 3; ++k) {
            im = RFMA(IM(T, i, k), RE(Z, j, k), im);
            im = RFMA(-RE(T, i, k), IM(Z, j, k), im);
        }
        RE(A, i, j) = re;
        IM(A, i, j) = im;
    }
}
/* A += T Z */
static void F(acc_nn)(REAL *A, const REAL *T, const REAL *Z) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
        REAL re = RE(A, i, j), im = IM(A, i, j);
        for (int k = 0; k < 3; ++k) {
            re = RFMA(RE(T, i, k), RE(Z, k, j), re);
            re = RFMA(-IM(T, i, k), IM(Z, k, j), re);
        }
        for (int k = 0; k < 3; ++k) {
            im = RFMA(RE(T, i, k), IM(Z, k, j), im);
            im = RFMA(IM(T, i, k), RE(Z, k, j), im);
        }
        RE(A, i, j) = re;
        IM(A, i, j) = im;
    }
}

/* FP32 staple from the first two rows of each neighbour (the device reads
 * rows 0 and 1 only): row 2 = conj(row 0 x row 1), rebuilt where a product
 * needs it.  Entry (2, c) = conj(a_j b_k - a_k b_j), (c, j, k) cyclic. */
static void F(recon_row2)(REAL *M) {
    for (int c = 0; c < 3; ++c) {
        int j = (c + 1) % 3, k = (c + 2) % 3;
        REAL re = RE(M, 0, j) * RE(M, 1, k);
        re = RFMA(-IM(M, 0, j), IM(M, 1, k), re);
        re = RFMA(-RE(M, 0, k), RE(M, 1, j), re);
        re = RFMA(IM(M, 0, k), IM(M, 1, j), re);
        REAL im = RE(M, 0, j) * IM(M, 1, k);
        im = RFMA(IM(M, 0, j), RE(M, 1, k), im);
        im = RFMA(-RE(M, 0, k), IM(M, 1, j), im);
        im = RFMA(-IM(M, 0, k), RE(M, 1, j), im);
        RE(M, 2, c) = re;
        IM(M, 2, c) = -im;
    }
}
/* rows 0 and 1 of T = X Y */
static void F(mul_nn_rows01)(const REAL *X, const REAL *Y, REAL *T) {
    for (int i = 0; i < 2; ++i) for (int j = 0; j < 3; ++j) {
        REAL re = RE(X, i, 0) * RE(Y, 0, j);
        re = RFMA(-IM(X, i, 0), IM(Y, 0, j), re);
        re = RFMA(RE(X, i, 1), RE(Y, 1, j), re);
        re = RFMA(-IM(X, i, 1), IM(Y, 1, j), re);
        re = RFMA(RE(X, i, 2), RE(Y, 2, j), re);
        re = RFMA(-IM(X, i, 2), IM(Y, 2, j), re);
        REAL im = RE(X, i, 0) * IM(Y, 0, j);
        im = RFMA(IM(X, i, 0), RE(Y, 0, j), im);
        im = RFMA(RE(X, i, 1), IM(Y, 1, j), im);
        im = RFMA(IM(X, i, 1), RE(Y, 1, j), im);
        im = RFMA(RE(X, i, 2), IM(Y, 2, j), im);
        im = RFMA(IM(X, i, 2), RE(Y, 2, j), im);
        RE(T, i, j) = re;
        IM(T, i, j) = im;
    }
}
/* rows 0 and 1 of T = X Y^dagger (mul_nd's chain) */
static void F(mul_nd_rows01)(const REAL *X, const REAL *Y, REAL *T) {
    for (int i = 0; i < 2; ++i) for (int j = 0; j < 3; ++j) {
        REAL re = RE(X, i, 0) * RE(Y, j, 0);
        re = RFMA(IM(X, i, 0), IM(Y, j, 0), re);
        re = RFMA(RE(X, i, 1), RE(Y, j, 1), re);
        re = RFMA(IM(X, i, 1), IM(Y, j, 1), re);
        re = RFMA(RE(X, i, 2), RE(Y, j, 2), re);
        re = RFMA(IM(X, i, 2), IM(Y, j, 2), re);
        REAL im = IM(X, i, 0) * RE(Y, j, 0);
        im = RFMA(-RE(X, i, 0), IM(Y, j, 0), im);
        im = RFMA(IM(X, i, 1), RE(Y, j, 1), im);
        im = RFMA(-RE(X, i, 1), IM(Y, j, 1), im);
        im = RFMA(IM(X, i, 2), RE(Y, j, 2), im);
        im = RFMA(-RE(X, i, 2), IM(Y, j, 2), im);
        RE(T, i, j) = re;
        IM(T, i, j) = im;
    }
}
/* A += W^dagger Z */
static void F(acc_dn)(REAL *A, const REAL *W, const REAL *Z) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
        REAL re = RE(A, i, j), im = IM(A, i, j);
        for (int k = 0; k < 3; ++k) {
            re = RFMA(RE(W, k, i), RE(Z, k, j), re);
            re = RFMA(IM(W, k, i), IM(Z, k, j), re);
        }
        for (int k = 0; k < 3; ++k) {
            im = RFMA(RE(W, k, i), IM(Z, k, j), im);
            im = RFMA(-IM(W, k, i), RE(Z, k, j), im);
        }
        RE(A, i, j) = re;
        IM(A, i, j) = im;
    }
}

/* A = sum_{nu != mu} [U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)].
 * FP64 storage: full products (mul_nd / acc_nd forward, mul_dd / acc_nn
 * backward).  FP32 storage: forward X (Z Y)^+ with Y, V = Z Y and P = X V^+
 * completed by recon_row2, A += P; backward (Y X)^+ Z with X, W = Y X and Z
 * completed, A += W^+ Z. */
static void F(su3_staple)(const REAL *L, const geom *g, const int64_t x[4], int mu, REAL *A) {
    for (int c = 0; c < 18; ++c) A[c] = 0;
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        int64_t xpm[4], xpn[4], xmn[4], xpmmn[4];
        shift(g, x, mu, +1, xpm);
        shift(g, x, nu, +1, xpn);
        shift(g, x, nu, -1, xmn);
        shift(g, xpm, nu, -1, xpmmn);
        REAL X[18], Y[18], Z[18], T[18];
        if (sizeof(REAL) == 4) {
            REAL V[18], P[18];
            F(load_link)(L, g, xpm, nu, 18, X);
            F(load_link)(L, g, xpn, mu, 18, Y);
            F(load_link)(L, g, x, nu, 18, Z);
            F(recon_row2)(Y);
            F(mul_nn_rows01)(Z, Y, V);
            F(recon_row2)(V);
            F(mul_nd_rows01)(X, V, P);
            F(recon_row2)(P);
            for (int c = 0; c < 18; ++c) A[c] = A[c] + P[c];
            F(load_link)(L, g, xpmmn, nu, 18, X);
            F(load_link)(L, g, xmn, mu, 18, Y);
            F(load_link)(L, g, xmn, nu, 18, Z);
            F(recon_row2)(X);
            F(mul_nn_rows01)(Y, X, T);
            F(recon_row2)(T);
            F(recon_row2)(Z);
            F(acc_dn)(A, T, Z);
            continue;
        }
        F(load_link)(L, g, xpm, nu, 18, X);
        F(load_link)(L, g, xpn, mu, 18, Y);
        F(mul_nd)(X, Y, T);
        F(load_link)(L, g, x, nu, 18, Z);
        F(acc_nd)(A, T, Z);
        F(load_link)(L, g, xpmmn, nu, 18, X);
        F(load_link)(L, g, xmn, mu, 18, Y);
        F(mul_dd)(X, Y, T);
        F(load_link)(L, g, xmn, nu, 18, Z);
        F(acc_nn)(A, T, Z);
    }
}
static int F(su3_staple_out)(const int64_t n[4], const void *dev, const int64_t x[4], int mu, double *out18) {
    geom g = make_geom(n);
    REAL A[18];
    F(su3_staple)((const REAL *)dev, &g, x, mu, A);
    for (int c = 0; c < 18; ++c) out18[c] = (double)A[c];
    return YMO_OK;
}

/* ---------------------------------------------------- FP64 link algebra */
#ifndef YM_ROT_DEFINED
#define YM_ROT_DEFINED
/* rows i,j <- [[c, i s],[i s, c]] rows (lambda_1-type block) */
static inline void rot_sym(double *m, int i, int j, double c, double s) {
    for (int k = 0; k < 3; ++k) {
        double xr = RE(m, i, k), xi = IM(m, i, k), yr = RE(m, j, k), yi = IM(m, j, k);
        RE(m, i, k) = fma(c, xr, -(s * yi));
        IM(m, i, k) = fma(c, xi, s * yr);
        RE(m, j, k) = fma(c, yr, -(s * xi));
        IM(m, j, k) = fma(c, yi, s * xr);
    }
}
/* rows i,j <- [[c, s],[-s, c]] rows (lambda_2-type block) */
static inline void rot_asym(double *m, int i, int j, double c, double s) {
    for (int k = 0; k < 3; ++k) {
        double xr = RE(m, i, k), xi = IM(m, i, k), yr = RE(m, j, k), yi = IM(m, j, k);
        RE(m, i, k) = fma(c, xr, s * yr);
        IM(m, i, k) = fma(c, xi, s * yi);
        RE(m, j, k) = fma(c, yr, -(s * xr));
        IM(m, j, k) = fma(c, yi, -(s * xi));
    }
}
/* Parts of M = u a used by a proposal (a the staple in FP64), as single
   FMA chains over k: re_chain continues acc with +-Re (u a)_{rb}, im_chain
   with +-Im (u a)_{rb}; `first` starts the chain with a product, `neg`
   subtracts.  Each proposal reads the rows of the current u. */
static inline double re_chain(const double *u, const double *a, int r, int b, double acc, int first, int neg) {
    for (int k = 0; k < 3; ++k) {
        double ur = neg ? -RE(u, r, k) : RE(u, r, k);
        double ui = neg ? IM(u, r, k) : -IM(u, r, k);
        acc = (first && k == 0) ? ur * RE(a, k, b) : fma(ur, RE(a, k, b), acc);
        acc = fma(ui, IM(a, k, b), acc);
    }
    return acc;
}
static inline double im_chain(const double *u, const double *a, int r, int b, double acc, int first, int neg) {
    for (int k = 0; k < 3; ++k) {
        double ur = neg ? -RE(u, r, k) : RE(u, r, k);
        double ui = neg ? -IM(u, r, k) : IM(u, r, k);
        acc = (first && k == 0) ? ur * IM(a, k, b) : fma(ur, IM(a, k, b), acc);
        acc = fma(ui, RE(a, k, b), acc);
    }
    return acc;
}
/* row i <- (c + i s) row i */
static inline void rot_phase(double *m, int i, double c, double s) {
    for (int k = 0; k < 3; ++k) {
        double xr = RE(m, i, k), xi = IM(m, i, k);
        RE(m, i, k) = fma(c, xr, -(s * xi));
        IM(m, i, k) = fma(c, xi, s * xr);
    }
}
#endif

#ifndef YM_REUNIT_DEFINED
#define YM_REUNIT_DEFINED
/* Gram-Schmidt on rows 0,1 and row2 = conj(row0 x row1), all FP64. */
static void reunitarize3(double *u) {
    double n = RE(u, 0, 0) * RE(u, 0, 0);
    n = fma(IM(u, 0, 0), IM(u, 0, 0), n);
    n = fma(RE(u, 0, 1), RE(u, 0, 1), n);
    n = fma(IM(u, 0, 1), IM(u, 0, 1), n);
    n = fma(RE(u, 0, 2), RE(u, 0, 2), n);
    n = fma(IM(u, 0, 2), IM(u, 0, 2), n);
    double inv = 1.0 / sqrt(n);
    for (int k = 0; k < 3; ++k) { RE(u, 0, k) *= inv; IM(u, 0, k) *= inv; }
    double pr = RE(u, 0, 0) * RE(u, 1, 0);
    pr = fma(IM(u, 0, 0), IM(u, 1, 0), pr);
    pr = fma(RE(u, 0, 1), RE(u, 1, 1), pr);
    pr = fma(IM(u, 0, 1), IM(u, 1, 1), pr);
    pr = fma(RE(u, 0, 2), RE(u, 1, 2), pr);
    pr = fma(IM(u, 0, 2), IM(u, 1, 2), pr);
    double pi = RE(u, 0, 0) * IM(u, 1, 0);
    pi = fma(-IM(u, 0, 0), RE(u, 1, 0), pi);
    pi = fma(RE(u, 0, 1), IM(u, 1, 1), pi);
    pi = fma(-IM(u, 0, 1), RE(u, 1, 1), pi);
    pi = fma(RE(u, 0, 2), IM(u, 1, 2), pi);
    pi = fma(-IM(u, 0, 2), RE(u, 1, 2), pi);
    for (int k = 0; k < 3; ++k) {
        double ar = RE(u, 0, k), ai = IM(u, 0, k);
        RE(u, 1, k) = fma(-pr, ar, fma(pi, ai, RE(u, 1, k)));
        IM(u, 1, k) = fma(-pr, ai, fma(-pi, ar, IM(u, 1, k)));
    }
    n = RE(u, 1, 0) * RE(u, 1, 0);
    n = fma(IM(u, 1, 0), IM(u, 1, 0), n);
    n = fma(RE(u, 1, 1), RE(u, 1, 1), n);
    n = fma(IM(u, 1, 1), IM(u, 1, 1), n);
    n = fma(RE(u, 1, 2), RE(u, 1, 2), n);
    n = fma(IM(u, 1, 2), IM(u, 1, 2), n);
    inv = 1.0 / sqrt(n);
    for (int k = 0; k < 3; ++k) { RE(u, 1, k) *= inv; IM(u, 1, k) *= inv; }
    /* row2_k = conj(a_p b_q - a_q b_p) for (k,p,q) cyclic */
    static const int cyc[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    for (int k = 0; k < 3; ++k) {
        int p = cyc[k][0], q = cyc[k][1];
        double ar = RE(u, 0, p), ai = IM(u, 0, p), br = RE(u, 1, q), bi = IM(u, 1, q);
        double cr = RE(u, 0, q), ci = IM(u, 0, q), dr = RE(u, 1, p), di = IM(u, 1, p);
        double re = ar * br;
        re = fma(-ai, bi, re);
        re = fma(-cr, dr, re);
        re = fma(ci, di, re);
        double im = ar * bi;
        im = fma(ai, br, im);
        im = fma(-cr, di, im);
        im = fma(-ci, dr, im);
        RE(u, 2, k) = re;
        IM(u, 2, k) = -im;
    }
}
#endif

/* ------------------------------------------------ SU(3) Metropolis sweep */
static void F(su3_update_link)(REAL *L, const geom *g, const int64_t x[4], int mu, const ymd_su3_params *P,
                               uint64_t *acc) {
    REAL A[18], Uf[18];
    F(su3_staple)(L, g, x, mu, A);
    F(load_link)(L, g, x, mu, 18, Uf);
    /* FP32 storage: the link's third row rebuilt from rows 0 and 1 */
    if (sizeof(REAL) == 4) F(recon_row2)(Uf);
    double u[18], a[18];
    for (int c = 0; c < 18; ++c) { u[c] = (double)Uf[c]; a[c] = (double)A[c]; }
    double nbt = -(P->beta / 3.0);
    uint32_t site = lex_of(g, x);
    static const int blk[8][2] = {{0, 1}, {0, 1}, {0, 1}, {0, 2}, {0, 2}, {1, 2}, {1, 2}, {0, 1}};
    for (int h = 0; h < P->hits; ++h) {
        uint32_t o[4];
        philox_draw(P->key, site, word1(mu, KIND_SU3_HIT, (uint32_t)h), P->sweep, o);
        double theta;
        ymo_proposal_angle_from_u64(raw64(o[0], o[1]), P->angle, &theta);
        double uacc = unit53(o[2], o[3]);
        int gen = h & 7;
        double s, cm1, s2 = 0.0, cm2 = 0.0, dre;
        int bi = blk[gen][0], bj = blk[gen][1];
        /* t1 = Re (u a)_ii + Re (u a)_jj as one FMA chain */
        double t1 = re_chain(u, a, bj, bj, re_chain(u, a, bi, bi, 0.0, 1, 0), 0, 0);
        if (gen == 7) {
            double psi = theta * 0x1.279a74590331cp-2;
            sincos_hit_angle(P->angle, psi, &s, &cm1);
            sincos_hit_angle(P->angle, -2.0 * psi, &s2, &cm2);
            double t2 = im_chain(u, a, 1, 1, im_chain(u, a, 0, 0, 0.0, 1, 0), 0, 0);
            double r22 = re_chain(u, a, 2, 2, 0.0, 1, 0), i22 = im_chain(u, a, 2, 2, 0.0, 1, 0);
            double d1 = fma(cm1, t1, -(s * t2));
            double d2 = fma(cm2, r22, -(s2 * i22));
            dre = d1 + d2;
        } else {
            sincos_hit_angle(P->angle, theta * 0.5, &s, &cm1);
            if (gen == 2) {
                double t2 = im_chain(u, a, 1, 1, im_chain(u, a, 0, 0, 0.0, 1, 0), 0, 1);
                dre = fma(cm1, t1, -(s * t2));
            } else if (gen == 0 || gen == 3 || gen == 5) {
                double t2 = im_chain(u, a, bj, bi, im_chain(u, a, bi, bj, 0.0, 1, 0), 0, 0);
                dre = fma(cm1, t1, -(s * t2));
            } else {
                double t2 = re_chain(u, a, bi, bj, re_chain(u, a, bj, bi, 0.0, 1, 0), 0, 1);
                dre = fma(cm1, t1, s * t2);
            }
        }
        double dS = nbt * dre;
        int accept = (dS <= 0.0) || (uacc < ymd_exp_neg(dS));
        *acc += accept ? 1 : 0;
        /* branch-free: a rejected proposal applies (c, s) = (1, 0) */
        double c = accept ? 1.0 + cm1 : 1.0;
        double sa = accept ? s : 0.0;
        if (gen == 7) {
            double c2 = accept ? 1.0 + cm2 : 1.0;
            double s2a = accept ? s2 : 0.0;
            rot_phase(u, 0, c, sa); rot_phase(u, 1, c, sa); rot_phase(u, 2, c2, s2a);
        } else if (gen == 2) {
            rot_phase(u, 0, c, sa); rot_phase(u, 1, c, -sa);
        } else if (gen == 0 || gen == 3 || gen == 5) {
            rot_sym(u, bi, bj, c, sa);
        } else {
            rot_asym(u, bi, bj, c, sa);
        }
    }
    if (P->reunitarize) reunitarize3(u);
    for (int c = 0; c < 18; ++c) Uf[c] = (REAL)u[c];
    F(store_link)(L, g, x, mu, 18, Uf);
}

static int F(su3_sweep)(const int64_t n[4], void *dev, const ymd_su3_params *P, uint64_t *proposals, uint64_t *accepts) {
    geom g = make_geom(n);
    if (g.n[0] % 2 || g.n[1] % 2 || g.n[2] % 2 || g.n[3] % 2) return YMO_UNSUPPORTED;
    REAL *L = (REAL *)dev;
    uint64_t acc = 0, prop = 0;
    for (int mu = 0; mu < 4; ++mu)
        for (int p = 0; p < 2; ++p) {
            /* the links of one sub-pass are independent: any split over
             * threads gives the same links and the same integer counts */
            YMD_PARALLEL_FOR_REDUCE(+ : acc, prop)
            for (int64_t s = 0; s < g.vh; ++s) {
                int64_t x[4];
                if (!site_of(&g, p, s, x)) continue;
                F(su3_update_link)(L, &g, x, mu, P, &acc);
                prop += (uint64_t)P->hits;
            }
        }
    *proposals = prop;
    *accepts = acc;
    return YMO_OK;
}

/* ---------------------------------------------------------- plaquette */
/* Average plaquette (1/6V) sum_x sum_{mu<nu} (1/3) Re Tr P_{mu nu}(x) of the
 * stored links, products in FP64; one partial sum per x0 slice (threads over
 * slices), the slices added in order, so the value does not depend on the
 * thread count.  A measurement for CPU chains in statistical comparisons,
 * not a replay of a device kernel. */
static void F(m3d)(const REAL *L, const geom *g, const int64_t x[4], int mu, double m[18]) {
    REAL v[18];
    F(load_link)(L, g, x, mu, 18, v);
    for (int c = 0; c < 18; ++c) m[c] = (double)v[c];
}
/* r = a b (dagger_b: a b^dagger) */
static void F(m3mul)(const double *a, const double *b, int dagger_b, double *r) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 3; ++k) {
                double ar = a[(3 * i + k) * 2], ai = a[(3 * i + k) * 2 + 1];
                double br = dagger_b ? b[(3 * j + k) * 2] : b[(3 * k + j) * 2];
                double bi = dagger_b ? -b[(3 * j + k) * 2 + 1] : b[(3 * k + j) * 2 + 1];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            r[(3 * i + j) * 2] = re;
            r[(3 * i + j) * 2 + 1] = im;
        }
}
static int F(su3_plaq)(const int64_t n[4], const void *dev, double *out) {
    geom g = make_geom(n);
    const REAL *L = (const REAL *)dev;
    double *part = (double *)calloc((size_t)n[0], sizeof(double));
    if (!part) return YMO_ALLOCATION_FAILED;
    YMD_PARALLEL_FOR
    for (int64_t x0 = 0; x0 < n[0]; ++x0) {
        double acc = 0.0;
        for (int64_t x1 = 0; x1 < n[1]; ++x1) for (int64_t x2 = 0; x2 < n[2]; ++x2) for (int64_t x3 = 0; x3 < n[3]; ++x3) {
            const int64_t x[4] = {x0, x1, x2, x3};
            for (int mu = 0; mu < 4; ++mu)
                for (int nu = mu + 1; nu < 4; ++nu) {
                    int64_t xm[4], xn[4];
                    shift(&g, x, mu, +1, xm);
                    shift(&g, x, nu, +1, xn);
                    double a[18], b[18], c[18], d[18], t1[18], t2[18], t3[18];
                    F(m3d)(L, &g, x, mu, a);
                    F(m3d)(L, &g, xm, nu, b);
                    F(m3d)(L, &g, xn, mu, c);
                    F(m3d)(L, &g, x, nu, d);
                    F(m3mul)(a, b, 0, t1);
                    F(m3mul)(t1, c, 1, t2);
                    F(m3mul)(t2, d, 1, t3);
                    acc += (t3[0] + t3[8] + t3[16]) / 3.0;
                }
        }
        part[x0] = acc;
    }
    double tot = 0.0;
    for (int64_t x0 = 0; x0 < n[0]; ++x0) tot += part[x0];
    free(part);
    *out = tot / (6.0 * (double)(n[0] * n[1] * n[2] * n[3]));
    return YMO_OK;
}

/* ---------------------------------------------------------- hot starts */
/* Box-Muller from two 53-bit uniforms: r = sqrt(-2 log(1-u1)), angle 2 pi u2 */
static inline void F(gauss2)(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1, double *g1, double *g2) {
    double u1 = 1.0 - unit53(a0, a1);
    double u2 = unit53(b0, b1);
    double r = sqrt(-2.0 * ymd_log(u1));
    double s, c;
    ymd_sincos_turns(u2, &s, &c);
    *g1 = r * c;
    *g2 = r * s;
}
static int F(su3_hot)(const int64_t n[4], void *dev, uint64_t key) {
    geom g = make_geom(n);
    REAL *L = (REAL *)dev;
    YMD_PARALLEL_FOR
    for (int64_t x0 = 0; x0 < n[0]; ++x0)
    for (int64_t x1 = 0; x1 < n[1]; ++x1) for (int64_t x2 = 0; x2 < n[2]; ++x2) for (int64_t x3 = 0; x3 < n[3]; ++x3) {
        const int64_t x[4] = {x0, x1, x2, x3};
        uint32_t site = lex_of(&g, x);
        for (int mu = 0; mu < 4; ++mu) {
            double gs[12];
            for (int k = 0; k < 6; ++k) {
                uint32_t o[4];
                philox_draw(key, site, word1(mu, KIND_HOT_START, (uint32_t)k), 0, o);
                F(gauss2)(o[0], o[1], o[2], o[3], &gs[2 * k], &gs[2 * k + 1]);
            }
            double u[18];
            for (int c = 0; c < 12; ++c) u[c] = gs[c];
            for (int c = 12; c < 18; ++c) u[c] = 0.0;
            reunitarize3(u);
            REAL v[18];
            for (int c = 0; c < 18; ++c) v[c] = (REAL)u[c];
            F(store_link)(L, &g, x, mu, 18, v);
        }
    }
    return YMO_OK;
}

/* ---------------------------------------------------------- SU(2) */
/* quaternion product with the reference convention (lib.rs:212-224) as an
 * explicit FMA chain */
static inline void F(qmul)(const REAL *a, const REAL *b, REAL *r) {
    REAL r0 = a[0] * b[0];
    r0 = RFMA(-a[1], b[1], r0); r0 = RFMA(-a[2], b[2], r0); r0 = RFMA(-a[3], b[3], r0);
    REAL r1 = a[0] * b[1];
    r1 = RFMA(a[1], b[0], r1); r1 = RFMA(a[2], b[3], r1); r1 = RFMA(-a[3], b[2], r1);
    REAL r2 = a[0] * b[2];
    r2 = RFMA(-a[1], b[3], r2); r2 = RFMA(a[2], b[0], r2); r2 = RFMA(a[3], b[1], r2);
    REAL r3 = a[0] * b[3];
    r3 = RFMA(a[1], b[2], r3); r3 = RFMA(-a[2], b[1], r3); r3 = RFMA(a[3], b[0], r3);
    r[0] = r0; r[1] = r1; r[2] = r2; r[3] = r3;
}
static inline void F(qdag)(const REAL *a, REAL *r) { r[0] = a[0]; r[1] = -a[1]; r[2] = -a[2]; r[3] = -a[3]; }

static void F(su2_staple)(const REAL *L, const geom *g, const int64_t x[4], int mu, REAL *A) {
    for (int c = 0; c < 4; ++c) A[c] = 0;
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        int64_t xpm[4], xpn[4], xmn[4], xpmmn[4];
        shift(g, x, mu, +1, xpm);
        shift(g, x, nu, +1, xpn);
        shift(g, x, nu, -1, xmn);
        shift(g, xpm, nu, -1, xpmmn);
        REAL X[4], Y[4], Z[4], Yd[4], Zd[4], Xd[4], T[4], R[4];
        F(load_link)(L, g, xpm, nu, 4, X);
        F(load_link)(L, g, xpn, mu, 4, Y);
        F(load_link)(L, g, x, nu, 4, Z);
        F(qdag)(Y, Yd); F(qdag)(Z, Zd);
        F(qmul)(X, Yd, T);
        F(qmul)(T, Zd, R);
        for (int c = 0; c < 4; ++c) A[c] = A[c] + R[c];
        F(load_link)(L, g, xpmmn, nu, 4, X);
        F(load_link)(L, g, xmn, mu, 4, Y);
        F(load_link)(L, g, xmn, nu, 4, Z);
        F(qdag)(X, Xd); F(qdag)(Y, Yd);
        F(qmul)(Xd, Yd, T);
        F(qmul)(T, Z, R);
        for (int c = 0; c < 4; ++c) A[c] = A[c] + R[c];
    }
}


static int F(su2_hb)(const int64_t n[4], void *dev, double beta, uint64_t key, uint64_t sweep, int64_t max_attempts,
                     uint64_t *proposals, uint64_t *updates) {
    geom g = make_geom(n);
    if (g.n[0] % 2 || g.n[1] % 2 || g.n[2] % 2 || g.n[3] % 2) return YMO_UNSUPPORTED;
    REAL *L = (REAL *)dev;
    uint64_t prop = 0, upd = 0;
    int status = YMO_OK;
    int failed = 0;
    for (int mu = 0; mu < 4; ++mu)
        for (int p = 0; p < 2; ++p) {
            YMD_PARALLEL_FOR_REDUCE(+ : prop, upd, failed)
            for (int64_t s = 0; s < g.vh; ++s) {
                int64_t x[4];
                if (!site_of(&g, p, s, x)) continue;
                REAL A[4];
                F(su2_staple)(L, &g, x, mu, A);
                double a[4] = {(double)A[0], (double)A[1], (double)A[2], (double)A[3]};
                double un[4];
                int64_t att = 0;
                int ok = ymd_su2_heatbath_link(a, beta, key, lex_of(&g, x), mu, sweep, max_attempts, un, &att);
                prop += (uint64_t)att;
                if (!ok) { failed += 1; continue; }
                upd += 1;
                REAL v[4] = {(REAL)un[0], (REAL)un[1], (REAL)un[2], (REAL)un[3]};
                F(store_link)(L, &g, x, mu, 4, v);
            }
        }
    if (failed) status = YMO_HEATBATH_ATTEMPT_LIMIT;
    *proposals = prop;
    *updates = upd;
    return status;
}

static int F(su2_or)(const int64_t n[4], void *dev) {
    geom g = make_geom(n);
    if (g.n[0] % 2 || g.n[1] % 2 || g.n[2] % 2 || g.n[3] % 2) return YMO_UNSUPPORTED;
    REAL *L = (REAL *)dev;
    for (int mu = 0; mu < 4; ++mu)
        for (int p = 0; p < 2; ++p) {
            YMD_PARALLEL_FOR
            for (int64_t s = 0; s < g.vh; ++s) {
                int64_t x[4];
                if (!site_of(&g, p, s, x)) continue;
                REAL A[4], Uf[4];
                F(su2_staple)(L, &g, x, mu, A);
                double a[4] = {(double)A[0], (double)A[1], (double)A[2], (double)A[3]};
                double norm = sqrt(qnorm2(a));
                if (norm == 0.0) continue;
                double ad[4] = {a[0] / norm, -(a[1] / norm), -(a[2] / norm), -(a[3] / norm)};
                F(load_link)(L, &g, x, mu, 4, Uf);
                double ud[4] = {(double)Uf[0], -(double)Uf[1], -(double)Uf[2], -(double)Uf[3]};
                double t[4], r[4];
                qmul64(ad, ud, t);
                qmul64(t, ad, r);
                qnormalize(r);
                REAL v[4] = {(REAL)r[0], (REAL)r[1], (REAL)r[2], (REAL)r[3]};
                F(store_link)(L, &g, x, mu, 4, v);
            }
        }
    return YMO_OK;
}

static int F(su2_hot)(const int64_t n[4], void *dev, uint64_t key) {
    geom g = make_geom(n);
    REAL *L = (REAL *)dev;
    YMD_PARALLEL_FOR
    for (int64_t x0 = 0; x0 < n[0]; ++x0)
    for (int64_t x1 = 0; x1 < n[1]; ++x1) for (int64_t x2 = 0; x2 < n[2]; ++x2) for (int64_t x3 = 0; x3 < n[3]; ++x3) {
        const int64_t x[4] = {x0, x1, x2, x3};
        uint32_t site = lex_of(&g, x);
        for (int mu = 0; mu < 4; ++mu) {
            double q[4];
            for (int k = 0; k < 2; ++k) {
                uint32_t o[4];
                philox_draw(key, site, word1(mu, KIND_HOT_START, (uint32_t)k), 0, o);
                F(gauss2)(o[0], o[1], o[2], o[3], &q[2 * k], &q[2 * k + 1]);
            }
            qnormalize(q);
            REAL v[4] = {(REAL)q[0], (REAL)q[1], (REAL)q[2], (REAL)q[3]};
            F(store_link)(L, &g, x, mu, 4, v);
        }
    }
    return YMO_OK;
}

#undef RE
#undef IM
#undef F
#undef YM_CAT
#undef YM_CAT2
