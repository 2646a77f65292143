/*
 * bse_c_driver.c -- a plain C caller of libbse_mi355x.so (no Python, no
 * torch): reads the reference's input files, runs the q = 0 TDA path and
 * writes the results as text.
 *
 *   bse_c_driver <params> <out>
 *
 * <params> is a key/value file that stands for the constants the reference's
 * main() sets in code (bse.version13.cpp:3686-3843); its paths name the
 * reference's own input files: the Wannier90 hr.dat (one block, or two
 * blocks up/down for collinear spin; Hamiltonian_TB ctor, 460-560) and the
 * centres .xyz (589-601).  Keys (one per line, '#' starts a comment):
 *   hr_dat <path>            hr_blocks <1|2>
 *   centres <path>           centres_atoms <number of atom lines per block>
 *   lattice a1x a1y a1z a2x a2y a2z a3x a3y a3z    (Angstrom, rows a_i)
 *   grid n1 n2 n3            shift s1 s2 s3        (in grid steps)
 *   bands n_occ nv nc        scissor x             q qx qy qz
 *   screening r0 eps_env     g_cut x               spin_factor f_s
 *   scales exchange direct   dipole_dir x y z      dimension 2|3
 *   radius x                 precision 0|1|2       n_vectors k (-1: all)
 *   refine 0|1               eta x                 omega w0 w1 n_omega
 *   screening_model model|rpa                      rpa_window nv nc (0 0: all)
 * <out> gets "N <n>", then "E <energy> <|f|^2>" per exciton, then
 * "eps <omega> <re> <im>" per frequency, then "Z <k>" and the first k
 * eigenvectors (reference row order) as "re im" lines, column by column.
 * Exit status 0 on success; every library error is printed with its text.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bse_mi355x.h"

#define CHECK(call)                                                                  \
    do {                                                                             \
        int rc_ = (call);                                                            \
        if (rc_ != BSE_OK) {                                                         \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_,                      \
                    ctx ? bse_last_error(ctx) : "no context");                       \
            return 1;                                                                \
        }                                                                            \
    } while (0)

typedef struct {
    char hr_dat[1024], centres[1024];
    int hr_blocks, centres_atoms;
    double lattice[9]; /* rows a_i */
    int grid[3];
    double shift[3];
    int bands[3];
    double scissor, q[3], r0, eps_env, g_cut, f_s, scales[2], dipole_dir[3];
    int dimension;
    double radius;
    int precision, n_vectors, refine;
    double eta, omega[2];
    int n_omega;
    int screening, rpa_nv, rpa_nc;
} params;

static void join_dir(char* dst, size_t cap, const char* base_file, const char* rel) {
    if (rel[0] == '/') {
        snprintf(dst, cap, "%s", rel);
        return;
    }
    const char* slash = strrchr(base_file, '/');
    if (!slash) {
        snprintf(dst, cap, "%s", rel);
        return;
    }
    snprintf(dst, cap, "%.*s/%s", (int)(slash - base_file), base_file, rel);
}

static int read_params(const char* path, params* p) {
    FILE* f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        return 1;
    }
    memset(p, 0, sizeof(*p));
    p->hr_blocks = 1;
    p->grid[2] = 1;
    p->f_s = 2.0;
    p->scales[0] = p->scales[1] = 1.0;
    p->dipole_dir[0] = 1.0;
    p->dimension = 2;
    p->n_vectors = -1;
    char line[4096], key[64], val[1024];
    while (fgets(line, sizeof line, f)) {
        char* hash = strchr(line, '#');
        if (hash) *hash = 0;
        if (sscanf(line, "%63s", key) != 1) continue;
        const char* rest = line + strspn(line, " \t");
        rest += strlen(key);
        int ok = 1;
        if (!strcmp(key, "hr_dat")) {
            ok = sscanf(rest, "%1023s", val) == 1;
            if (ok) join_dir(p->hr_dat, sizeof p->hr_dat, path, val);
        } else if (!strcmp(key, "centres")) {
            ok = sscanf(rest, "%1023s", val) == 1;
            if (ok) join_dir(p->centres, sizeof p->centres, path, val);
        } else if (!strcmp(key, "hr_blocks")) ok = sscanf(rest, "%d", &p->hr_blocks) == 1;
        else if (!strcmp(key, "centres_atoms")) ok = sscanf(rest, "%d", &p->centres_atoms) == 1;
        else if (!strcmp(key, "lattice"))
            ok = sscanf(rest, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", &p->lattice[0], &p->lattice[1],
                        &p->lattice[2], &p->lattice[3], &p->lattice[4], &p->lattice[5], &p->lattice[6],
                        &p->lattice[7], &p->lattice[8]) == 9;
        else if (!strcmp(key, "grid")) ok = sscanf(rest, "%d %d %d", &p->grid[0], &p->grid[1], &p->grid[2]) == 3;
        else if (!strcmp(key, "shift"))
            ok = sscanf(rest, "%lf %lf %lf", &p->shift[0], &p->shift[1], &p->shift[2]) == 3;
        else if (!strcmp(key, "bands")) ok = sscanf(rest, "%d %d %d", &p->bands[0], &p->bands[1], &p->bands[2]) == 3;
        else if (!strcmp(key, "scissor")) ok = sscanf(rest, "%lf", &p->scissor) == 1;
        else if (!strcmp(key, "q")) ok = sscanf(rest, "%lf %lf %lf", &p->q[0], &p->q[1], &p->q[2]) == 3;
        else if (!strcmp(key, "screening")) ok = sscanf(rest, "%lf %lf", &p->r0, &p->eps_env) == 2;
        else if (!strcmp(key, "g_cut")) ok = sscanf(rest, "%lf", &p->g_cut) == 1;
        else if (!strcmp(key, "spin_factor")) ok = sscanf(rest, "%lf", &p->f_s) == 1;
        else if (!strcmp(key, "scales")) ok = sscanf(rest, "%lf %lf", &p->scales[0], &p->scales[1]) == 2;
        else if (!strcmp(key, "dipole_dir"))
            ok = sscanf(rest, "%lf %lf %lf", &p->dipole_dir[0], &p->dipole_dir[1], &p->dipole_dir[2]) == 3;
        else if (!strcmp(key, "dimension")) ok = sscanf(rest, "%d", &p->dimension) == 1;
        else if (!strcmp(key, "radius")) ok = sscanf(rest, "%lf", &p->radius) == 1;
        else if (!strcmp(key, "precision")) ok = sscanf(rest, "%d", &p->precision) == 1;
        else if (!strcmp(key, "n_vectors")) ok = sscanf(rest, "%d", &p->n_vectors) == 1;
        else if (!strcmp(key, "refine")) ok = sscanf(rest, "%d", &p->refine) == 1;
        else if (!strcmp(key, "eta")) ok = sscanf(rest, "%lf", &p->eta) == 1;
        else if (!strcmp(key, "omega")) ok = sscanf(rest, "%lf %lf %d", &p->omega[0], &p->omega[1], &p->n_omega) == 3;
        else if (!strcmp(key, "screening_model")) {
            ok = sscanf(rest, "%1023s", val) == 1 && (!strcmp(val, "model") || !strcmp(val, "rpa"));
            p->screening = ok && !strcmp(val, "rpa");
        } else if (!strcmp(key, "rpa_window")) ok = sscanf(rest, "%d %d", &p->rpa_nv, &p->rpa_nc) == 2;
        else {
            fprintf(stderr, "%s: unknown key '%s'\n", path, key);
            ok = 0;
        }
        if (!ok) {
            fprintf(stderr, "%s: bad line for '%s'\n", path, key);
            fclose(f);
            return 1;
        }
    }
    fclose(f);
    if (!p->hr_dat[0] || !p->centres[0]) {
        fprintf(stderr, "%s: hr_dat and centres are required\n", path);
        return 1;
    }
    return 0;
}

/* One Wannier90 hr block: header line, num_wann, nrpts, nrpts ndegen values,
 * then nrpts*num_wann^2 lines "R1 R2 R3 l m re im" (read at 494-552).
 * H(R) is divided by ndegen (the Wannier90 meaning of the weights). */
static int read_hr_block(FILE* f, int* norb, int* nr, int** rint, double** h) {
    char line[4096];
    if (!fgets(line, sizeof line, f)) return 1;
    if (fscanf(f, "%d %d", norb, nr) != 2 || *norb <= 0 || *nr <= 0) return 2;
    int* deg = (int*)malloc(sizeof(int) * (size_t)*nr);
    for (int i = 0; i < *nr; ++i)
        if (fscanf(f, "%d", &deg[i]) != 1 || deg[i] <= 0) {
            free(deg);
            return 2;
        }
    const int n = *norb;
    *rint = (int*)calloc((size_t)*nr * 3, sizeof(int));
    *h = (double*)calloc((size_t)*nr * n * n * 2, sizeof(double));
    for (int ir = 0; ir < *nr; ++ir)
        for (int e = 0; e < n * n; ++e) {
            int r1, r2, r3, l, m;
            double re, im;
            if (fscanf(f, "%d %d %d %d %d %lf %lf", &r1, &r2, &r3, &l, &m, &re, &im) != 7 || l < 1 || l > n ||
                m < 1 || m > n) {
                free(deg);
                return 2;
            }
            (*rint)[3 * ir] = r1;
            (*rint)[3 * ir + 1] = r2;
            (*rint)[3 * ir + 2] = r3;
            const size_t at = (((size_t)ir * n + (l - 1)) * n + (m - 1)) * 2;
            (*h)[at] = re / deg[ir];
            (*h)[at + 1] = im / deg[ir];
        }
    free(deg);
    /* consume the rest of the last line so a second block starts cleanly */
    if (!fgets(line, sizeof line, f)) line[0] = 0;
    return 0;
}

/* Centres: count line, comment line, "X x y z" per Wannier function, then
 * the atom lines (skipped). */
static int read_centres_block(FILE* f, int norb, int atoms, double* tau, int stride, int off) {
    char line[4096];
    if (!fgets(line, sizeof line, f) || !fgets(line, sizeof line, f)) return 1;
    for (int i = 0; i < norb; ++i) {
        char sym[16];
        if (!fgets(line, sizeof line, f)) return 1;
        double* t = tau + 3 * (size_t)(stride * i + off);
        if (sscanf(line, "%15s %lf %lf %lf", sym, &t[0], &t[1], &t[2]) != 4) return 1;
    }
    for (int i = 0; i < atoms; ++i)
        if (!fgets(line, sizeof line, f)) return 1;
    return 0;
}

int main(int argc, char** argv) {
    bse_ctx* ctx = NULL;
    if (argc != 3) {
        fprintf(stderr, "usage: %s <params> <out>\n", argv[0]);
        return 2;
    }
    params p;
    if (read_params(argv[1], &p)) return 1;

    /* model: one block, or two collinear blocks merged to spinor order 2a+s */
    FILE* fh = fopen(p.hr_dat, "r");
    if (!fh) {
        fprintf(stderr, "cannot open %s\n", p.hr_dat);
        return 1;
    }
    int norb_b = 0, nr = 0, *rint = NULL;
    double* hb = NULL;
    if (read_hr_block(fh, &norb_b, &nr, &rint, &hb)) {
        fprintf(stderr, "%s: malformed hr block\n", p.hr_dat);
        return 1;
    }
    int norb = norb_b;
    double* h = hb;
    if (p.hr_blocks == 2) {
        int norb_d = 0, nr_d = 0, *rint_d = NULL;
        double* hd = NULL;
        if (read_hr_block(fh, &norb_d, &nr_d, &rint_d, &hd) || norb_d != norb_b || nr_d != nr ||
            memcmp(rint_d, rint, sizeof(int) * 3 * (size_t)nr)) {
            fprintf(stderr, "%s: the two spin blocks differ\n", p.hr_dat);
            return 1;
        }
        norb = 2 * norb_b;
        h = (double*)calloc((size_t)nr * norb * norb * 2, sizeof(double));
        for (int ir = 0; ir < nr; ++ir)
            for (int a = 0; a < norb_b; ++a)
                for (int b = 0; b < norb_b; ++b)
                    for (int s = 0; s < 2; ++s) {
                        const double* src = (s ? hd : hb) + (((size_t)ir * norb_b + a) * norb_b + b) * 2;
                        double* dst = h + (((size_t)ir * norb + 2 * a + s) * norb + 2 * b + s) * 2;
                        dst[0] = src[0];
                        dst[1] = src[1];
                    }
        free(rint_d);
        free(hd);
        free(hb);
    }
    fclose(fh);

    double* tau = (double*)calloc((size_t)norb * 3, sizeof(double));
    FILE* fc = fopen(p.centres, "r");
    if (!fc) {
        fprintf(stderr, "cannot open %s\n", p.centres);
        return 1;
    }
    for (int s = 0; s < p.hr_blocks; ++s)
        if (read_centres_block(fc, norb_b, p.centres_atoms, tau, p.hr_blocks, s)) {
            fprintf(stderr, "%s: malformed centres block\n", p.centres);
            return 1;
        }
    fclose(fc);

    /* Cartesian R = R1 a1 + R2 a2 + R3 a3 (Hamiltonian_TB::FFT, 647-705) */
    double* r_cart = (double*)calloc((size_t)nr * 3, sizeof(double));
    for (int ir = 0; ir < nr; ++ir)
        for (int c = 0; c < 3; ++c)
            r_cart[3 * ir + c] = rint[3 * ir] * p.lattice[c] + rint[3 * ir + 1] * p.lattice[3 + c] +
                                 rint[3 * ir + 2] * p.lattice[6 + c];

    bse_tb_model model = {norb, nr, r_cart, h, tau};
    bse_system sys;
    memset(&sys, 0, sizeof sys);
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) sys.bravais[3 * i + c] = p.lattice[3 * i + c]; /* column i = a_i */
    sys.n1 = p.grid[0];
    sys.n2 = p.grid[1];
    sys.n3 = p.grid[2];
    sys.shift[0] = p.shift[0];
    sys.shift[1] = p.shift[1];
    sys.shift3 = p.shift[2];
    sys.n_occ = p.bands[0];
    sys.nv = p.bands[1];
    sys.nc = p.bands[2];
    sys.scissor = p.scissor;
    memcpy(sys.q, p.q, sizeof p.q);
    sys.r0 = p.r0;
    sys.eps_env = p.eps_env;
    sys.g_cut = p.g_cut;
    sys.f_s = p.f_s;
    sys.exchange_scale = p.scales[0];
    sys.direct_scale = p.scales[1];
    memcpy(sys.dipole_dir, p.dipole_dir, sizeof p.dipole_dir);
    sys.dimension = p.dimension;
    sys.radius = p.radius;
    sys.screening = p.screening;
    sys.rpa_nv = p.rpa_nv;
    sys.rpa_nc = p.rpa_nc;

    if (bse_abi_version() != BSE_ABI_VERSION) {
        fprintf(stderr, "library ABI %d, header %d\n", bse_abi_version(), BSE_ABI_VERSION);
        return 1;
    }
    if (bse_create(&ctx, 0) != BSE_OK) {
        fprintf(stderr, "bse_create failed\n");
        return 1;
    }
    CHECK(bse_set_problem(ctx, &model, &sys));
    int64_t sizes[4];
    CHECK(bse_problem_sizes(ctx, sizes));
    const int n = (int)sizes[1];
    CHECK(bse_build(ctx, p.precision));
    const int nvec = p.n_vectors < 0 ? n : p.n_vectors;
    CHECK(bse_eigh(ctx, p.n_vectors, p.refine));

    double* E = (double*)malloc(sizeof(double) * (size_t)n);
    double* fosc = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* omega = (double*)malloc(sizeof(double) * (size_t)(p.n_omega > 0 ? p.n_omega : 1));
    double* eps_re = (double*)malloc(sizeof(double) * (size_t)(p.n_omega > 0 ? p.n_omega : 1));
    double* eps_im = (double*)malloc(sizeof(double) * (size_t)(p.n_omega > 0 ? p.n_omega : 1));
    double* z = (double*)malloc(sizeof(double) * 2 * (size_t)n * (nvec > 0 ? nvec : 1));
    CHECK(bse_get_eigenvalues(ctx, E));
    CHECK(bse_oscillator(ctx, fosc));
    for (int s = 0; s < p.n_omega; ++s)
        omega[s] = p.n_omega > 1 ? p.omega[0] + (p.omega[1] - p.omega[0]) * s / (p.n_omega - 1) : p.omega[0];
    if (p.n_omega > 0) CHECK(bse_spectrum(ctx, omega, p.n_omega, p.eta, eps_re, eps_im));
    if (nvec > 0) CHECK(bse_get_eigenvectors(ctx, z, n, nvec, 0));

    FILE* fo = fopen(argv[2], "w");
    if (!fo) {
        fprintf(stderr, "cannot write %s\n", argv[2]);
        return 1;
    }
    fprintf(fo, "N %d\n", n);
    for (int m = 0; m < n; ++m)
        fprintf(fo, "E %.17g %.17g\n", E[m], fosc[2 * m] * fosc[2 * m] + fosc[2 * m + 1] * fosc[2 * m + 1]);
    for (int s = 0; s < p.n_omega; ++s) fprintf(fo, "eps %.17g %.17g %.17g\n", omega[s], eps_re[s], eps_im[s]);
    fprintf(fo, "Z %d\n", nvec);
    for (size_t i = 0; i < (size_t)n * nvec; ++i) fprintf(fo, "%.17g %.17g\n", z[2 * i], z[2 * i + 1]);
    fclose(fo);
    CHECK(bse_destroy(ctx));
    printf("N = %d, E0 = %.12f eV\n", n, E[0]);
    free(E);
    free(fosc);
    free(omega);
    free(eps_re);
    free(eps_im);
    free(z);
    free(r_cart);
    free(tau);
    free(rint);
    free(h);
    return 0;
}
