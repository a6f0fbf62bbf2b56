/*
 * lem_oracle.c — CPU oracle for the PARALLEM per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see lem_oracle.h).  Every function restates the
 * reference function cited beside it; where the reference is racy, undefined
 * or schedule-dependent, the canonical choice below is used and named:
 *   D1  plateau routing in synchronous (Jacobi) rounds
 *   D2  init/merge of watershed labels bounds-checked (no phantom columns)
 *   D3  sink labels = 1 + rank of the sink cell in row-major order
 *   D4  exact edge set, per-pair minimum, no overflow; Kruskal edges in
 *       (weight, lower label, higher label) order, node1 = lower label
 *   D7  Slopes keep the pass-1 (unraised) values; pass-2 writes are scratch
 *   D9  sediment blocking = least fixed point of the reference's ok/localOK rounds
 *   D10 cells on an fd cycle are never computed (fa = w, sediment untouched)
 *   D11 a neighbour outside the grid behaves exactly like a nodata neighbour
 *   D12 FP64 transcendentals from include/lem_math.h, -ffp-contract=off
 *   D20 temperature/rain read from the same per-cell arrays as the device
 * Fixed-point loops (CalcMFD, MFDsedbud) are evaluated in topological order
 * with the reference's per-cell summation order; the *_literal variants
 * replay the reference's round structure and are used by the tests to show
 * the two agree.
 */
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "lem_oracle.h"
#include "../include/lem_math.h"

static const int DX[8] = {0, 1, 1, 1, 0, -1, -1, -1};   /* routeflow.cu:701 */
static const int DY[8] = {-1, -1, 0, 1, 1, 1, 0, -1};   /* routeflow.cu:702 */
static const int CODE[8] = {64, 128, 1, 2, 4, 8, 16, 32}; /* routeflow.h:32 */
/* pull order of CalcMFD / MFDsedbud (mfd_simple.cu:95-165, mfd_accum.cu:93-221):
 * neighbour positions E, SE, S, SW, W, NW, N, NE as direction indices */
static const int PULL[8] = {2, 3, 4, 5, 6, 7, 0, 1};

#define NODATA (-9999)

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ init */

/* memory.cu:217-242 */
void orc_createmask(int N, const double *dem, int *mask)
{
    for (int x = 0; x < N; x++) mask[x] = (dem[x] == -9999) ? 0 : 1;
}

/* Data.cpp:38-66: first strict minimum over mask cells, lowest starts at 1000 */
int orc_init_setoutletcell(int W, int H, const int *mask, const double *dem)
{
    double lowest = 1000.;
    int index = 0;
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            int cell = i * W + j;
            if (mask[cell] == 1 && dem[cell] < lowest) { lowest = dem[cell]; index = cell; }
        }
    return index;
}

/* memory.cu:197-214 */
void orc_createSoilTfromformula(int N, const double *dem, double *soilT)
{
    for (int c = 0; c < N; c++) soilT[c] = (dem[c] > 0) ? ((dem[c] - 400) / 1400) * 5 : 0.0;
}

/* Data.cpp:246-271: rain gradient x2 down the rows, lapse-rate temperature */
void orc_SetClimateMatrices(const orc_params *p, const int *mask, const double *dem,
                            double rain_scalar, double temp_scalar, double *rainmat, double *tempmat)
{
    int W = p->W, H = p->H;
    double rainmax = rain_scalar * 2;
    double rainmin = rain_scalar;
    double increment = (rainmax - rainmin) / H;
    double minheight = 467.228;
    for (int i = 0; i < H; i++) {
        double therain = rainmin + (i * increment);
        for (int j = 0; j < W; j++) {
            int c = i * W + j;
            if (mask[c] == 1) {
                rainmat[c] = therain;
                tempmat[c] = temp_scalar - ((dem[c] - minheight) * 0.0065);
            } else {
                rainmat[c] = -9999.;
                tempmat[c] = -9999.;
            }
        }
    }
}

/* Data.cpp:115-185 (values that feed the path; soilT comes from the formula) */
void orc_setProcessMatrices(const orc_params *p, orc_state *s)
{
    int N = p->W * p->H;
    double rescale = (10000 / (p->cellsize * p->cellsize));
    for (int x = 0; x < N; x++) {
        if (s->mask[x] == 1) {
            s->fd[x] = 0; s->SFD[x] = 0; s->fa[x] = 0.0; s->SlopePtr[x] = 0.0; s->dz[x] = 0.0;
            s->w[x] = 1.0;
            s->stone[x] = p->stoneval; s->fines[x] = p->fineval; s->soilM[x] = p->soilMval;
            s->nut[x] = p->nutval; s->soilB[x] = p->soilBval;
            double tempeffect = 100 + (100 / (s->temp[x] + 15));
            s->TotB[x] = (3000.0 + (5.0 * (s->rain[x] - tempeffect))) / rescale;
        } else {
            s->fd[x] = 0; s->SFD[x] = 0; s->fa[x] = 0.0; s->dz[x] = 0.0; s->SlopePtr[x] = 0.0;
            s->w[x] = 0.0; s->soilT[x] = 0.0; s->stone[x] = 0.0; s->fines[x] = 0.0;
            s->soilM[x] = 0.0; s->nut[x] = 0.0; s->soilB[x] = 0.0; s->TotB[x] = 0.0;
        }
        s->ero[x] = 0.0; s->geli[x] = 0.0; s->inci[x] = 0.0; s->depo[x] = 0.0;
        s->wC[x] = 0.0; s->wP[x] = 0.0;
    }
    for (long y = 0; y < 8L * N; y++) s->Slopes[y] = 0.0;
}

/* -------------------------------------------------------------------- FD */

/* FlowDirs, newflow.cu:49-174 */
void orc_flowdirs(int W, int H, int cell_size, const double *zs, const int *mask,
                  double *slopes, int *SFD, int *mfd, int *sinkcounter, int *flatcounter)
{
    int sinkc = 0, flatc = 0;
#pragma omp parallel for reduction(+ : sinkc, flatc) schedule(static)
    for (int irow = 0; irow < H; irow++) {
        for (int icol = 0; icol < W; icol++) {
            long self = (long)irow * W + icol;
            mfd[self] = -9999;
            SFD[self] = -9999;
            if (mask[self] != 1) continue;
            double smax = 0.0;
            int aspect = 0, aspectsfd = 0, upslopecount = 0, flatcount = 0;
            for (int d = 0; d < 8; d++) {
                int cellx = icol + DX[d], celly = irow + DY[d];
                if (cellx >= 0 && cellx < W && celly >= 0 && celly < H) {
                    float dc = (d % 2 == 0) ? 1.0f : 1.41f;
                    float den = (float)cell_size * dc;   /* int * float evaluated in float */
                    double z = zs[self];
                    double zn = zs[(long)celly * W + cellx];
                    double stemp = (z - zn) / (double)den;
                    if (zn == -9999) stemp = -0.0000001;
                    if ((z > zn) && (zn != -9999)) {
                        slopes[8 * self + d] = stemp;
                        if (stemp > smax) { smax = stemp; aspectsfd = CODE[d]; }
                        aspect = aspect + CODE[d];
                    }
                    if (z < zn) upslopecount += 1;
                    if (z == zn) flatcount += 1;
                }
            }
            if (upslopecount > 7) { sinkc++; aspect = 0; aspectsfd = 0; }
            if ((flatcount + upslopecount) > 7) {
                if (upslopecount <= 7) { flatc++; aspect = 0; aspectsfd = 0; }
            }
            if ((flatcount < 8) && (flatcount + upslopecount > 7)) { aspect = 0; aspectsfd = 0; }
            mfd[self] = aspect;
            SFD[self] = aspectsfd;
        }
    }
    if (sinkcounter) *sinkcounter += sinkc;
    if (flatcounter) *flatcounter += flatc;
}

/* shortest_paths_plateaus_initMFD (newflow.cu:234-256) followed by
 * route_plateausMFD (newflow.cu:264-350) looped as in routeflow.cu:854-872,
 * in synchronous rounds (D1).  Only cells whose initial distance is nonzero
 * can ever change, so each round scans that candidate list.  Returns the
 * number of rounds that changed something. */
int orc_route_plateaus(int W, int H, const double *zs, const int *mask,
                       double *slopes, int *fd, int *SFD, float *sp)
{
    long N = (long)W * H;
    long ncand = 0;
    for (long i = 0; i < N; i++) {
        if (mask[i] != 1) { sp[i] = 0.0f; continue; }
        sp[i] = 20000000;
        if (fd[i] > 0) sp[i] = 0;
        if (SFD[i] > 0) sp[i] = 0;
        if (sp[i] != 0) ncand++;
    }
    long *cand = (long *)malloc(sizeof(long) * (ncand + 1));
    float *prev = (float *)malloc(sizeof(float) * N);
    long k = 0;
    for (long i = 0; i < N; i++)
        if (mask[i] == 1 && sp[i] != 0) cand[k++] = i;
    int rounds = 0;
    int changed;
    /* only candidates ever change, so each round copies only them into the
     * previous-round buffer; within a round every candidate reads only that
     * buffer and writes only its own cell (Jacobi), so the candidates can be
     * processed in parallel */
    memcpy(prev, sp, sizeof(float) * N);
    do {
        changed = 0;
#pragma omp parallel for schedule(static)
        for (long c = 0; c < ncand; c++) prev[cand[c]] = sp[cand[c]];
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : changed)
        for (long c = 0; c < ncand; c++) {
            long self = cand[c];
            int irow = (int)(self / W), icol = (int)(self % W);
            float min_distance = prev[self];
            if (min_distance == 0) continue;
            for (int d = 0; d < 8; d++) {
                int cellx = icol + DX[d], celly = irow + DY[d];
                if (cellx >= 0 && cellx < W && celly >= 0 && celly < H) {
                    float dc = (d % 2 == 0) ? 1.0f : 1.41f;
                    long n = (long)celly * W + cellx;
                    float dis = prev[n] + dc;
                    if (zs[self] == zs[n] && min_distance > dis) {
                        min_distance = dis;
                        changed = 1;
                        fd[self] = CODE[d];
                        SFD[self] = CODE[d];
                        slopes[8 * self + d] = 0.0001;
                        sp[self] = min_distance;
                    }
                }
            }
        }
        if (changed) rounds++;
    } while (changed);
    free(cand);
    free(prev);
    return rounds;
}

/* The same rounds over every cell of the grid in row-major order (no candidate list). */
int orc_route_plateaus_literal(int W, int H, const double *zs, const int *mask,
                               double *slopes, int *fd, int *SFD, float *sp)
{
    long N = (long)W * H;
    for (long i = 0; i < N; i++) {
        if (mask[i] != 1) { sp[i] = 0.0f; continue; }
        sp[i] = 20000000;
        if (fd[i] > 0) sp[i] = 0;
        if (SFD[i] > 0) sp[i] = 0;
    }
    float *prev = (float *)malloc(sizeof(float) * N);
    int rounds = 0, changed;
    do {
        changed = 0;
        memcpy(prev, sp, sizeof(float) * N);
        for (int irow = 0; irow < H; irow++)
            for (int icol = 0; icol < W; icol++) {
                long self = (long)irow * W + icol;
                if (mask[self] != 1) continue;
                float min_distance = prev[self];
                if (min_distance == 0) continue;
                for (int d = 0; d < 8; d++) {
                    int cellx = icol + DX[d], celly = irow + DY[d];
                    if (cellx >= 0 && cellx < W && celly >= 0 && celly < H) {
                        float dc = (d % 2 == 0) ? 1.0f : 1.41f;
                        long n = (long)celly * W + cellx;
                        float dis = prev[n] + dc;
                        if (zs[self] == zs[n] && min_distance > dis) {
                            min_distance = dis; changed = 1;
                            fd[self] = CODE[d]; SFD[self] = CODE[d];
                            slopes[8 * self + d] = 0.0001;
                            sp[self] = min_distance;
                        }
                    }
                }
            }
        if (changed) rounds++;
    } while (changed);
    free(prev);
    return rounds;
}

static int code_to_dir(int code)
{
    for (int d = 0; d < 8; d++)
        if (CODE[d] == code) return d;
    return -1;
}

static long uf_find(long *par, long x)
{
    while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; }
    return x;
}

/* init_watershedMFD (newflow.cu:352-420, D2/D3), init_watershed_sinkMFD loop
 * (newflow.cu:430-460, routeflow.cu:285-302) and identify_watershedMFD loop
 * (newflow.cu:469-487, routeflow.cu:314-338).  Returns the number of sink
 * labels issued (labels are 1..nlab, 0 = border/nodata). */
int orc_watershed_labels(int W, int H, const double *zs, const int *mask, const int *SFD, int *ws)
{
    long N = (long)W * H;
    int counter = 0;
    /* sinks get fresh labels, other cells point at their SFD receiver */
    for (int irow = 0; irow < H; irow++)
        for (int icol = 0; icol < W; icol++) {
            long self = (long)irow * W + icol;
            if (mask[self] != 1) { ws[self] = 0; continue; }
            if ((icol == W - 1) || (irow == 0) || (irow == H - 1)) { ws[self] = 0; continue; }
            if (SFD[self] == 0) ws[self] = ++counter;
            else {
                int d = code_to_dir(SFD[self]);
                int cellx = icol + DX[d], celly = irow + DY[d];
                ws[self] = -(celly * W + cellx + 1);
            }
        }
    /* every mask cell with SFD==0 adopts the largest label among its
     * equal-height neighbours until nothing changes.  The fixed point gives
     * each 8-connected equal-height component of SFD==0 mask cells its
     * largest label; union-find reaches the same fixed point directly. */
    long *par = (long *)malloc(sizeof(long) * N);
    for (long i = 0; i < N; i++) par[i] = i;
    for (int irow = 0; irow < H; irow++)
        for (int icol = 0; icol < W; icol++) {
            long self = (long)irow * W + icol;
            if (mask[self] != 1 || SFD[self] != 0) continue;
            for (int d = 0; d < 8; d++) {
                int cellx = icol + DX[d], celly = irow + DY[d];
                if (cellx >= 0 && cellx < W && celly >= 0 && celly < H) {
                    long n = (long)celly * W + cellx;
                    if (mask[n] == 1 && SFD[n] == 0 && zs[n] == zs[self]) {
                        long a = uf_find(par, self), b = uf_find(par, n);
                        if (a != b) par[a] = b;
                    }
                }
            }
        }
    int *cmax = (int *)malloc(sizeof(int) * N);
    for (long i = 0; i < N; i++) cmax[i] = -1;
    for (long i = 0; i < N; i++)
        if (mask[i] == 1 && SFD[i] == 0) {
            long r = uf_find(par, i);
            if (ws[i] > cmax[r]) cmax[r] = ws[i];
        }
    for (long i = 0; i < N; i++)
        if (mask[i] == 1 && SFD[i] == 0) ws[i] = cmax[uf_find(par, i)];
    free(cmax);
    free(par);
    /* pointer jumping to the terminal label (chains are acyclic: SFD
     * pointers strictly decrease height or, on plateaus, distance) */
    long *stack = (long *)malloc(sizeof(long) * N);
    for (long i = 0; i < N; i++) {
        if (mask[i] != 1 || ws[i] >= 0) continue;
        long top = 0, c = i;
        while (ws[c] < 0) { stack[top++] = c; c = (long)(-ws[c] - 1); }
        int lab = ws[c];
        while (top > 0) ws[stack[--top]] = lab;
    }
    free(stack);
    return counter;
}

typedef struct { uint64_t key; double w; } orc_edge;

static int cmp_key_w(const void *a, const void *b)
{
    const orc_edge *x = (const orc_edge *)a, *y = (const orc_edge *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->w < y->w) return -1;
    if (x->w > y->w) return 1;
    return 0;
}

static int cmp_w_key(const void *a, const void *b)
{
    const orc_edge *x = (const orc_edge *)a, *y = (const orc_edge *)b;
    if (x->w < y->w) return -1;
    if (x->w > y->w) return 1;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return 0;
}

/* identify_watershed_boundary_write_backMFD (newflow.cu:787-877, D4),
 * init_and_sort_watershed_bound_labelsMFD (newflow.cu:905-1050) and the
 * union-find pass of floodingDriver (routeflow.cu:510-572).  raise has nlab+1
 * entries.  Returns the number of distinct edges. */
int orc_spill_heights(int W, int H, const double *zs, const int *mask, const int *ws,
                      int nlab, double *raise, int *sinkcount, int *nedges)
{
    long N = (long)W * H;
    long cap = 1 << 20, ne = 0;
    orc_edge *e = (orc_edge *)malloc(sizeof(orc_edge) * cap);
    for (int irow = 0; irow < H; irow++)
        for (int icol = 0; icol < W; icol++) {
            long c = (long)irow * W + icol;
            if (mask[c] != 1) continue;
            int this_id = ws[c];
            for (int d = 0; d < 8; d++) {
                int cellx = icol + DX[d], celly = irow + DY[d];
                if (cellx >= 0 && cellx < W && celly >= 0 && celly < H) {
                    long n = (long)celly * W + cellx;
                    int nei_id = ws[n];
                    if (this_id != nei_id) {
                        if (!(zs[c] >= zs[n])) continue;
                        unsigned lo = (unsigned)(this_id < nei_id ? this_id : nei_id);
                        unsigned hi = (unsigned)(this_id < nei_id ? nei_id : this_id);
                        if (ne == cap) { cap *= 2; e = (orc_edge *)realloc(e, sizeof(orc_edge) * cap); }
                        e[ne].key = ((uint64_t)lo << 32) | hi;
                        e[ne].w = zs[c];
                        ne++;
                    }
                }
            }
        }
    (void)N;
    qsort(e, ne, sizeof(orc_edge), cmp_key_w);
    long m = 0;
    for (long i = 0; i < ne; i++)
        if (m == 0 || e[i].key != e[m - 1].key) e[m++] = e[i];   /* first = minimum weight */
    qsort(e, m, sizeof(orc_edge), cmp_w_key);
    /* distinct labels in the edge list (data->sinkcount, newflow.cu:1045) */
    char *seen = (char *)calloc(nlab + 1, 1);
    int nseen = 0;
    for (long i = 0; i < m; i++) {
        unsigned lo = (unsigned)(e[i].key >> 32), hi = (unsigned)(e[i].key & 0xffffffffu);
        if (!seen[lo]) { seen[lo] = 1; nseen++; }
        if (!seen[hi]) { seen[hi] = 1; nseen++; }
    }
    free(seen);
    /* union_set (newflow.cu:654-707) and Kruskal (routeflow.cu:529-572) */
    int *parent = (int *)calloc(nlab + 1, sizeof(int));
    int *rank = (int *)calloc(nlab + 1, sizeof(int));
    char *done = (char *)calloc(nlab + 1, 1);
    for (int i = 0; i <= nlab; i++) raise[i] = 0.0;
    done[0] = 1;
    for (long i = 0; i < m; i++) {
        int l1 = (int)(e[i].key >> 32), l2 = (int)(e[i].key & 0xffffffffu);
        double h = e[i].w;
        if (!(l1 == 0 || parent[l1] > 0)) { parent[l1] = l1; rank[l1] = 0; }
        if (!(l2 == 0 || parent[l2] > 0)) { parent[l2] = l2; rank[l2] = 0; }
        int r1 = l1; while (parent[r1] != r1) r1 = parent[r1];
        int r2 = l2; while (parent[r2] != r2) r2 = parent[r2];
        if (done[r1] && done[r2]) continue;
        if (done[r1] || done[r2]) {
            if (done[r1]) { done[r2] = 1; raise[r2] = h; }
            else { done[r1] = 1; raise[r1] = h; }
            continue;
        }
        if (r1 != r2) {
            if (rank[r1] > rank[r2]) parent[r2] = r1;
            else { parent[r1] = r2; if (rank[r1] == rank[r2]) rank[r2]++; }
            raise[r2] = raise[r1] = h;
        }
    }
    for (int i = 0; i <= nlab; i++)
        if (raise[i] != 0) {
            int r = i; while (parent[r] != r) r = parent[r];
            raise[i] = raise[r];
        }
    free(parent); free(rank); free(done); free(e);
    if (sinkcount) *sinkcount = nseen;
    if (nedges) *nedges = (int)m;
    return (int)m;
}

/* cuFlowDirection, routeflow.cu:686-975 (with floodingDriver :202-676) */
void orc_flow_direction(const orc_params *p, orc_state *s, int iter)
{
    (void)iter;
    int W = p->W, H = p->H;
    long N = (long)W * H;
    int csize = (int)p->cellsize;
    /* the device copy of dem is taken before the outlet check (:731 vs :741) */
    double *zfd = (double *)malloc(sizeof(double) * N);
    memcpy(zfd, s->dem, sizeof(double) * N);
    double lowest = 1000.;                        /* testoutletislowest, :171-184 */
    for (long x = 0; x < N; x++)
        if (s->mask[x] == 1 && s->dem[x] < lowest) lowest = s->dem[x];
    if (lowest < s->dem[s->outlet]) s->dem[s->outlet] = lowest - 0.0001;

    int sinkc = 0, flatc = 0;
    orc_flowdirs(W, H, csize, zfd, s->mask, s->Slopes, s->SFD, s->fd, &sinkc, &flatc);
    float *sp = (float *)malloc(sizeof(float) * N);
    s->plateau_rounds1 = orc_route_plateaus(W, H, zfd, s->mask, s->Slopes, s->fd, s->SFD, sp);
    s->plateau_rounds2 = 0;
    s->nlabels = 0;
    s->nedges = 0;
    int sinkcount = 0;
    for (long j = 0; j < N; j++)
        if (s->mask[j] == 1 && s->SFD[j] == 0) sinkcount++;
    s->sinkcount = sinkcount;
    s->sinks_pass1 = sinkcount;
    if (sinkcount > 0) {
        int *ws = (int *)malloc(sizeof(int) * N);
        int nlab = orc_watershed_labels(W, H, zfd, s->mask, s->SFD, ws);
        double *raise = (double *)malloc(sizeof(double) * (nlab + 1));
        int sc = 0, nedges = 0;
        orc_spill_heights(W, H, zfd, s->mask, ws, nlab, raise, &sc, &nedges);
        for (long i = 0; i < N; i++)                  /* raise_watershedMFD, newflow.cu:879-894 */
            if (s->mask[i] == 1 && zfd[i] < raise[ws[i]]) zfd[i] = raise[ws[i]];
        double *scratch = (double *)calloc(8 * N, sizeof(double));
        int a = 0, b = 0;
        orc_flowdirs(W, H, csize, zfd, s->mask, scratch, s->SFD, s->fd, &a, &b);
        s->plateau_rounds2 = orc_route_plateaus(W, H, zfd, s->mask, scratch, s->fd, s->SFD, sp);
        s->sinkcount = sc;
        s->nlabels = nlab;
        s->nedges = nedges;
        free(scratch); free(raise); free(ws);
    }
    if (p->flowtype == 0)
        for (long i = 0; i < N; i++) s->fd[i] = s->SFD[i];
    if (p->flowtype == 1)
        for (long i = 0; i < N; i++)
            if ((s->fd[i] <= 0) && (s->mask[i] == 1)) s->fd[i] = s->SFD[i];
    s->fd[s->outlet] = 16;
    free(sp);
    free(zfd);
}

/* -------------------------------------------------------------------- FA */

/* calcrunoffweight, runoffweight.cu:21-58 */
void orc_runoff(const orc_params *p, orc_state *s)
{
    long N = (long)p->W * p->H;
#pragma omp parallel for schedule(static)
    for (long self = 0; self < N; self++) {
        if (s->mask[self] == 0) { s->w[self] = -9999; continue; }
        double runoff_coeff = (1.0 - 0.0085 * (s->stone[self])) * lm_exp(-0.0025 * (s->TotB[self]) / 76.5)
                              * lm_tanh(0.6 + (s->soilM[self]) / 0.2);
        double myppt = s->rain[self] / 1000;
        s->w[self] = myppt * runoff_coeff;
        s->soilM[self] += myppt * (1. - runoff_coeff);
        if ((s->soilM[self]) > 0.5) s->soilM[self] = 0.5;
    }
    /* calcwater, runoffweight.cu:356-377: mask cells in index order */
    double sumwater = 0;
    for (long i = 0; i < N; i++)
        if (s->mask[i] == 1) sumwater += s->w[i];
    s->totwater = sumwater;
}

/* neighbour index in direction d or -1 outside the grid */
static long nbr(int W, int H, long i, int d)
{
    int r = (int)(i / W) + DY[d], c = (int)(i % W) + DX[d];
    if (r < 0 || r >= H || c < 0 || c >= W) return -1;
    return (long)r * W + c;
}

/* donor test of CalcMFD/MFDsedbud: the cell at direction d from i flows into i */
static int donor_bit(int d) { return CODE[(d + 4) & 7]; }

/* CalcMFD fixed point (mfd_simple.cu:33-171, loops :204-222, :293-305) in
 * topological order: a mask cell is computed once every real donor is, as
 * fa = w + max(0, sum of donor fa in the order E,SE,S,SW,W,NW,N,NE); nodata
 * (and, D11, off-grid) neighbours are ok from the start and add 0*0.  Cells
 * that never complete keep fa = w (D10). */
static int fa_topological(int W, int H, const int *mask, const int *fd, const double *w, double *fa)
{
    long N = (long)W * H;
    int *cnt = (int *)malloc(sizeof(int) * N);
    long *queue = (long *)malloc(sizeof(long) * N);
    long qh = 0, qt = 0;
    for (long i = 0; i < N; i++) {
        fa[i] = (mask[i] == 1) ? w[i] : 0.0;
        cnt[i] = 0;
        if (mask[i] != 1) continue;
        for (int k = 0; k < 8; k++) {
            int d = PULL[k];
            long n = nbr(W, H, i, d);
            if (n >= 0 && mask[n] == 1 && (fd[n] & donor_bit(d))) cnt[i]++;
        }
        if (cnt[i] == 0) queue[qt++] = i;
    }
    int done = 0;
    while (qh < qt) {
        long i = queue[qh++];
        double accum = 0.;
        for (int k = 0; k < 8; k++) {
            int d = PULL[k];
            long n = nbr(W, H, i, d);
            if (n >= 0 && mask[n] == 1 && (fd[n] & donor_bit(d))) accum += fa[n] * 1.0;
        }
        if (accum < 0.0) accum = 0.0;
        fa[i] = w[i];
        fa[i] += accum;
        done++;
        for (int d = 0; d < 8; d++) {
            if (!(fd[i] & CODE[d])) continue;
            long r = nbr(W, H, i, d);
            if (r < 0 || mask[r] != 1) continue;
            if (--cnt[r] == 0) queue[qt++] = r;
        }
    }
    int active = 0;
    for (long i = 0; i < N; i++) active += (mask[i] == 1);
    free(cnt);
    free(queue);
    return active - done;
}

/* the reference's round structure replayed (one sequential sweep = one launch) */
int orc_flow_accum_literal(int W, int H, const int *mask, const int *fd, const double *w, double *fa)
{
    long N = (long)W * H;
    int *ok = (int *)malloc(sizeof(int) * N);
    int *localOK = (int *)malloc(sizeof(int) * N);
    for (long x = 0; x < N; x++) {
        fa[x] = 0.0;
        ok[x] = localOK[x] = (mask[x] == 1) ? 0 : 1;
    }
    int oneZero = 0, oneZero2;
    unsigned progress;
    do {
        progress = 0;
        oneZero2 = oneZero;
        for (long self = 0; self < N; self++) {
            if (mask[self] != 1) continue;
            if (ok[self] != 0) continue;
            if (localOK[self] == 1) { ok[self] = 1; continue; }
            double accum = 0.;
            fa[self] = w[self];
            int blocked = 0;
            for (int k = 0; k < 8 && !blocked; k++) {
                int d = PULL[k];
                long n = nbr(W, H, self, d);
                int nfd = (n >= 0) ? fd[n] : NODATA;
                if (nfd & donor_bit(d)) {
                    int nok = (n >= 0) ? ok[n] : 1;
                    if (!nok) { blocked = 1; break; }
                    double nfa = (n >= 0) ? fa[n] : 0.0;
                    double prop = (n >= 0 && mask[n] == 1) ? 1.0 : 0.0;
                    accum += nfa * prop;
                }
            }
            if (blocked) continue;
            if (accum < 0.0) accum = 0.0;
            fa[self] += accum;
            localOK[self] = 1;
            progress++;
        }
        if (progress == 0) oneZero = 1;
        else { oneZero = 0; oneZero2 = 0; }
    } while ((progress > 5) || (!oneZero2));
    int uncomputed = 0;
    for (long x = 0; x < N; x++) uncomputed += (mask[x] == 1 && ok[x] != 1);
    free(ok);
    free(localOK);
    return uncomputed;
}

/* correctmfdflow, mfd_simple.cu:238-384 */
void orc_flow_accum(const orc_params *p, orc_state *s, int iter)
{
    int W = p->W, H = p->H;
    long N = (long)W * H;
    s->fa_uncomputed = fa_topological(W, H, s->mask, s->fd, s->w, s->fa);
    int firstiter = p->start_iter - p->runiniter;
    if (iter == firstiter) {               /* :337-349 */
        double cpuFA_max = 0.0;
        int FAindex = 0;
        for (long i = 0; i < N; i++)
            if (s->fa[i] > cpuFA_max) { cpuFA_max = s->fa[i]; FAindex = (int)i; }
        s->FA_max = cpuFA_max;
        s->outlet = FAindex;
    } else {
        double m = 0.0;
        for (long i = 0; i < N; i++)
            if (s->fa[i] > m) m = s->fa[i];
        s->FA_max = m;
    }
}

/* ------------------------------------------------------------------- MOD */

/* MFDsedbud per-cell body once all donors are ok (mfd_accum.cu:63-252) */
static void sed_cell(int W, int H, long self, const int *mask, const int *fd,
                     const double *ero, const double *inci, const double *geli,
                     double ddk, double dck, double dgk,
                     double *depo, double *stone, double *fines, double *soilT)
{
    double accum_d = 0.0, accum_c = 0.0, accum_g = 0.0;
    double from[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; k++) {
        int d = PULL[k];
        long n = nbr(W, H, self, d);
        if (n >= 0 && mask[n] == 1 && (fd[n] & donor_bit(d))) {
            double prop = 1.0;
            accum_d += prop * (ero[n] * (1 - ddk));
            accum_c += prop * (inci[n] * (1 - dck));
            accum_g += prop * (geli[n] * (1 - dgk));
            from[k] = prop * stone[n];
        }
    }
    depo[self] = ((accum_d + ero[self]) * ddk) + ((accum_c + inci[self]) * dck) + ((accum_g + geli[self]) * dgk);
    double sedthick = (soilT[self]) + (depo[self]);
    if (sedthick != 0.0) {
        double old = (soilT[self]) / sedthick;
        double input = 1 - old;
        double propStonefrom = (from[0] + from[1] + from[2] + from[3] + from[4] + from[5] + from[6] + from[7]);
        stone[self] = (old * (stone[self])) + (input * (propStonefrom));
        if (stone[self] < 0.01) stone[self] = 0.01;
        else if (stone[self] > 99.9) stone[self] = 99.9;
        fines[self] = 100 - stone[self];
    }
    double toterosion = ero[self] + inci[self] + geli[self];
    if (toterosion >= sedthick) soilT[self] = 0.0;
    if (sedthick > toterosion) soilT[self] = sedthick - toterosion;
}

/* a neighbour position that can never become ok in MFDsedbud: nodata or
 * off-grid (fd = -9999 pattern) at E, SE, S, SW, W, or a real donor in row 0
 * or column 0 (mfd_accum.cu:42) */
static int sed_permanently_blocked(int W, int H, long i, const int *mask, const int *fd)
{
    for (int k = 0; k < 8; k++) {
        int d = PULL[k];
        long n = nbr(W, H, i, d);
        if (n < 0 || mask[n] != 1) {
            if (NODATA & donor_bit(d)) return 1;
            continue;
        }
        if (fd[n] & donor_bit(d)) {
            if ((n / W) == 0 || (n % W) == 0) return 1;
        }
    }
    return 0;
}

static int sed_topological(const orc_params *p, orc_state *s, double ddk, double dck, double dgk)
{
    int W = p->W, H = p->H;
    long N = (long)W * H;
    const int *mask = s->mask, *fd = s->fd;
    int *cnt = (int *)malloc(sizeof(int) * N);
    char *ready_ok = (char *)malloc(N);
    long *queue = (long *)malloc(sizeof(long) * N);
    long qh = 0, qt = 0;
    for (long i = 0; i < N; i++) {
        cnt[i] = 0;
        ready_ok[i] = 0;
        if (mask[i] != 1) continue;
        if ((i / W) == 0 || (i % W) == 0) continue;
        if (sed_permanently_blocked(W, H, i, mask, fd)) continue;
        ready_ok[i] = 1;
        for (int k = 0; k < 8; k++) {
            int d = PULL[k];
            long n = nbr(W, H, i, d);
            if (n >= 0 && mask[n] == 1 && (fd[n] & donor_bit(d))) cnt[i]++;
        }
        if (cnt[i] == 0) queue[qt++] = i;
    }
    int done = 0;
    while (qh < qt) {
        long i = queue[qh++];
        sed_cell(W, H, i, mask, fd, s->ero, s->inci, s->geli, ddk, dck, dgk, s->depo, s->stone, s->fines, s->soilT);
        done++;
        for (int d = 0; d < 8; d++) {
            if (!(fd[i] & CODE[d])) continue;
            long r = nbr(W, H, i, d);
            if (r < 0 || mask[r] != 1) continue;
            if (--cnt[r] == 0 && ready_ok[r]) queue[qt++] = r;
        }
    }
    int active = 0;
    for (long i = 0; i < N; i++) active += (mask[i] == 1);
    free(cnt); free(ready_ok); free(queue);
    return active - done;
}

/* the reference's sediment round structure replayed (one sweep = one launch);
 * expects depo already set to 1e-5 on mask cells; returns uncomputed cells */
int orc_sediment_literal(const orc_params *p, orc_state *s)
{
    int W = p->W, H = p->H;
    long N = (long)W * H;
    double ddk = p->ddk, dck = p->dck, dgk = p->dgk;
    if (ddk > 0.95) ddk = 0.95;
    if (dck > 0.9) dck = 0.9;
    int *ok = (int *)calloc(N, sizeof(int));
    int *localOK = (int *)calloc(N, sizeof(int));
    int oneZero = 0, oneZero2;
    unsigned progress;
    do {
        progress = 0;
        oneZero2 = oneZero;
        for (long self = 0; self < N; self++) {
            if ((self % W) == 0 || (self / W) == 0) continue;
            if (s->mask[self] == 0) continue;
            if (ok[self] != 0) continue;
            if (localOK[self] == 1) { ok[self] = 1; continue; }
            int blocked = 0;
            for (int k = 0; k < 8; k++) {
                int d = PULL[k];
                long n = nbr(W, H, self, d);
                int nfd = (n >= 0) ? s->fd[n] : NODATA;
                if (nfd & donor_bit(d)) {
                    int nok = (n >= 0) ? ok[n] : 0;
                    if (!nok) { blocked = 1; break; }
                }
            }
            if (blocked) continue;
            sed_cell(W, H, self, s->mask, s->fd, s->ero, s->inci, s->geli, ddk, dck, dgk,
                     s->depo, s->stone, s->fines, s->soilT);
            localOK[self] = 1;
            progress++;
        }
        if (progress == 0) oneZero = 1;
        else { oneZero = 0; oneZero2 = 0; }
    } while ((progress > 10 || !oneZero2));
    int unc = 0;
    for (long x = 0; x < N; x++) unc += (s->mask[x] == 1 && ok[x] != 1);
    free(ok); free(localOK);
    return unc;
}

/* erosionGPU, erosion.cu:69-186, with the kernels of eroincidep.cu,
 * mfd_accum.cu and updates.cu in the same order */
void orc_erosion(const orc_params *p, orc_state *s, int iter)
{
    (void)iter;
    int W = p->W, H = p->H;
    long N = (long)W * H;
    double cellsize = p->cellsize;
    /* aveslope, eroincidep.cu:8-37 */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) {
        long self = i * 8;
        s->SlopePtr[i] = 0.0;
        if (s->mask[i] == 1) {
            double total = 0;
            int nonzero = 0;
            for (int j = 0; j < 8; j++) {
                total += s->Slopes[self + j];
                if (s->Slopes[self + j] > 0) nonzero++;
            }
            if ((total > 0) && (nonzero > 0)) s->SlopePtr[i] = total / nonzero;
        }
    }
    /* diff_erosion :39-87, conc_erosion :89-152, gelifluction :154-198 */
    double rescale_c = 1 / (10000 / (cellsize * cellsize));
#pragma omp parallel for schedule(static)
    for (long self = 0; self < N; self++) {
        if (s->mask[self] == 0) { s->ero[self] = 0.0; s->inci[self] = 0.0; s->geli[self] = 0.0; continue; }
        /* diffuse */
        double erodibility = (2.77e-6 * lm_pow(s->fines[self], 1.14) * (12. - s->nut[self]))
                             + (0.043 * (p->soil_struct - 2.)) + (0.033 * (p->profile_permeability - 3.));
        if (erodibility < 0.0) erodibility = 0.0;
        double OF = ((s->fa[self]) * 1000) / 25;
        if (OF > p->Qthreshold) OF = p->Qthreshold;
        double slope = (s->SlopePtr[self]);
        if (slope == 0.0) slope = 0.0001;
        if (slope > 5.0) slope = 5.0;
        double answer = ((erodibility * OF * OF * lm_pow(slope, 1.67) *
                          (lm_pow((110. - (s->stone[self])), 0.877) / 61.7023))) *
                        lm_exp(-0.07 * (s->TotB[self]) / 76.5);
        if (answer > (s->soilT[self])) answer = s->soilT[self];
        s->ero[self] = answer;
        /* concentrated */
        double cslope = s->SlopePtr[self];
        if (cslope > 0.5) cslope = 0.5;
        double max_incision = p->max_incision;
        double f_temp = ((s->fa[self]) * rescale_c) - p->Qthreshold;
        double inci;
        if (f_temp < 0.0) { f_temp = 0.0; inci = 0.0; }
        else {
            if (cslope == 0.0)
                inci = p->Flow_A * lm_pow(f_temp, p->Flow_B) * 0.01 * lm_exp(-0.07 * (s->TotB[self]) / 76.5);
            else
                inci = p->Flow_A * lm_pow(f_temp, p->Flow_B) * (cslope) * lm_exp(-0.07 * (s->TotB[self]) / 76.5);
        }
        if (inci > max_incision) inci = max_incision;
        if (inci > s->soilT[self]) {
            double softrockinc = s->soilT[self];
            double bedrockinc = (1.0 - (softrockinc / inci)) * 0.0001 * lm_pow(f_temp, p->Flow_B) * (cslope)
                                * lm_exp(-0.07 * (s->TotB[self]) / 76.5);
            inci = softrockinc + bedrockinc;
            if (inci > max_incision) inci = max_incision;
        }
        s->inci[self] = inci;
        /* gelifluction */
        double kappa = p->kappa;
        double T = s->temp[self];
        if (T >= -20. && T <= 8.87)
            kappa = 0.15711 - 0.011835 * T - 0.001195 * T * T + 0.000038 * T * T * T + 0.000002 * T * T * T * T;
        double scaler = 1 - (s->fa[self] / p->Gflowlim);
        if (s->fa[self] > p->Gflowlim) scaler = 0.0;
        double slopethreshold = s->SlopePtr[self];
        if (slopethreshold > p->Gslope) slopethreshold = p->Gslope;
        double g = kappa * slopethreshold * scaler;
        if ((s->ero[self] + g) > s->soilT[self]) g = s->soilT[self] - s->ero[self];
        if (g < 0.0) g = 0.0;
        s->geli[self] = g;
    }
    double tE = 0, tI = 0, tG = 0;
    for (long i = 0; i < N; i++) { tE += s->ero[i]; tI += s->inci[i]; tG += s->geli[i]; }
    s->totE = tE; s->totI = tI; s->totG = tG;

    /* sedmfdaccum, mfd_accum.cu:345-483 */
    double ddk = p->ddk, dck = p->dck, dgk = p->dgk;
    if (ddk > 0.95) ddk = 0.95;
    if (dck > 0.9) dck = 0.9;
    for (long x = 0; x < N; x++)
        if (s->mask[x] == 1) s->depo[x] = 0.00001;
    s->sed_uncomputed = sed_topological(p, s, ddk, dck, dgk);
    double tD = 0;
    for (long i = 0; i < N; i++) tD += s->depo[i];
    s->totD = tD;

    /* surface_change :200-217 and weathering :220-327 (type 1, rescale :440) */
    double rescale_w = 10000 * (10000 / (cellsize * cellsize));
#pragma omp parallel for schedule(static)
    for (long self = 0; self < N; self++) {
        if (s->mask[self] == 0) { s->dz[self] = 0.0; s->wP[self] = 0.0; s->wC[self] = 0.0; continue; }
        s->dz[self] = (s->depo[self] - s->ero[self] - s->inci[self] - s->geli[self]);
        double T = s->temp[self], rain = s->rain[self];
        double PE_d = 1.44 * lm_pow(T + 6., 2);
        double P0_d = 7.7e-5, P1_d = 2.3;
        double t_abs = T + 273.16;
        double t2 = t_abs * t_abs;
        double ltemp = lm_log10(t_abs);
        double DHA = -0.4883 + 8.074e-4 * T;
        double DHB = -0.3241 + 1.600e-4 * T;
        double sqI = 0.31622776601683794; /* sqrt(0.1), correctly rounded */
        double denom1 = (-4. * DHA * sqI) / (1. + 5.e-8 * DHB * sqI);
        double gammCa_d = lm_pow(10., denom1);
        double gammHCO3_d = lm_pow(10., (-DHA * sqI) / (1. + 5.4e-8 * DHB * sqI));
        double K1_d = lm_pow(10., -356.3094 - 0.06091964 * t_abs + 21834.37 / t_abs + 126.8339 * ltemp - 1684915 / t2);
        double K2_d = lm_pow(10., -107.8871 - 0.03252849 * t_abs + 5151.79 / t_abs + 38.92561 * ltemp - 563713.9 / t2);
        double KC_d = lm_pow(10., -171.9065 - 0.077993 * t_abs + 2839.319 / t_abs + 71.595 * ltemp);
        double KH_d = lm_pow(10., 108.3865 + 0.01985076 * t_abs - 6919.53 / t_abs - 40.45154 * ltemp + 669365.0 / t2);
        double surface_flow = s->fa[self] * 1000;
        double evap_trans;
        if (rain + surface_flow <= PE_d) evap_trans = rain + surface_flow;
        else evap_trans = rain + surface_flow - PE_d;
        double PCO2;
        if (s->soilT[self] > 0.1) PCO2 = lm_pow(10., -2.5 + ((T - 15.) / 10.));
        else PCO2 = lm_pow(10., -3.5 + ((T - 15.) / 10.));
        double Ca2eq = lm_pow(PCO2 * ((K1_d * KC_d * KH_d) / (4. * K2_d * gammCa_d * gammHCO3_d * gammHCO3_d)), (1. / 3.));
        double wC = 0.0148148148 * Ca2eq * (surface_flow / rescale_w);
        if (wC <= 0.0) wC = 0.0;
        if (wC > 0.001) wC = 0.001;
        double wP = P0_d * lm_exp(-(P1_d) * s->soilT[self]);
        s->wC[self] = wC;
        s->wP[self] = wP;
        s->soilT[self] = s->soilT[self] + wC * 0.02 + wP;
        if (s->soilT[self] <= 0.0) {
            s->soilT[self] = 0.0; s->stone[self] = 0.0; s->fines[self] = 0.0; s->soilM[self] = 0.02;
        } else {
            if (wP > 0)
                s->stone[self] += s->stone[self] * (lm_tanh(((s->ero[self] / s->soilT[self]))) * -lm_tanh(((wC / wP))));
            else
                s->stone[self] += s->stone[self] * lm_tanh(((s->ero[self] / s->soilT[self])));
            if (s->stone[self] < 0.01) s->stone[self] = 0.01;
            else if (s->stone[self] > 99.9) s->stone[self] = 99.9;
            s->fines[self] = 100. - s->stone[self];
            s->soilM[self] -= (1.e-3 * evap_trans / s->soilT[self]);
            if (s->soilM[self] < 0.02) s->soilM[self] = 0.02;
            else if (s->soilM[self] > 0.5) s->soilM[self] = 0.5;
        }
    }
    double twc = 0, twp = 0;
    for (long i = 0; i < N; i++) { twc += s->wC[i]; twp += s->wP[i]; }
    s->totwC = twc; s->totwP = twp;

    /* update_newSurface, updates.cu:142-181 */
    double base_level = s->dem[s->outlet] - 0.00012;
#pragma omp parallel for schedule(static)
    for (long self = 0; self < N; self++) {
        if (s->mask[self] == 0) continue;
        s->dem[self] = s->dem[self] + s->dz[self] - s->wC[self] - s->wP[self];
        if (s->dem[self] <= base_level) s->dem[self] = base_level + 0.01;
    }
    s->dem[s->outlet] = base_level;
    s->dem[s->outlet] = s->dem[s->outlet] - 0.00012;

    /* update_nutrient_content, updates.cu:48-76 */
#pragma omp parallel for schedule(static)
    for (long self = 0; self < N; self++) {
        if (s->mask[self] == 0) continue;
        if (s->soilT[self] <= 0.0) s->nut[self] = 0.0;
        double leaf_fall = s->TotB[self] / (72.8 * lm_log(s->TotB[self] + 1.4));
        double decomposition = s->soilB[self] * lm_exp(0.12 * s->temp[self] - 5.75);
        s->soilB[self] = s->soilB[self] + leaf_fall - decomposition;
        if (s->soilT[self] == 0.0 && s->soilB[self] > 0.0) s->nut[self] = 100.0;
        else if (s->soilT[self] == 0.0) s->nut[self] = 0.0;
        else s->nut[self] = (s->soilB[self] / (1300. * s->soilT[self])) * 100.;
        if (s->nut[self] > 100) s->nut[self] = 100.;
    }

    /* grow_vegetation, updates.cu:78-139 (rescale = Data.cpp:124) */
    double rescale_v = (10000 / (cellsize * cellsize));
#pragma omp parallel for schedule(static)
    for (long self = 0; self < N; self++) {
        double T = s->temp[self];
        double delta_temp = T - s->last_temp[self];
        double tempeffect = 100 + (100 / (T + 15));
        double total_cc = (3000.0 + (5.0 * (s->rain[self] - tempeffect))) / rescale_v;
        if (s->mask[self] == 0) { s->TotB[self] = 0; continue; }
        if (delta_temp > 0) {
            double r_plus = delta_temp * (1.0 + lm_tanh((T - p->temp_inflex) / p->temp_sens));
            s->TotB[self] += ((r_plus * s->TotB[self]) * (1. - (s->TotB[self] / total_cc)));
        }
        if (delta_temp < 0) {
            double r_minus = delta_temp * (1.0 - lm_tanh((T - p->temp_inflex) / p->temp_sens));
            s->TotB[self] += ((r_minus * s->TotB[self]) * (s->TotB[self] / total_cc));
        }
        if (delta_temp == 0) s->TotB[self] = total_cc;
        if (s->TotB[self] < 1) s->TotB[self] = 1.0;
    }
    double tb = 0;
    for (long i = 0; i < N; i++) tb += s->TotB[i];
    s->totbio = tb;
}

/* lem.cpp:83-97 */
void orc_step(const orc_params *p, orc_state *s, int iter)
{
    orc_flow_direction(p, s, iter);
    orc_runoff(p, s);
    orc_flow_accum(p, s, iter);
    orc_erosion(p, s, iter);
}

/* fn: 0 exp, 1 log, 2 log10, 3 pow, 4 tanh, 5 expm1 */
void orc_math_eval(int fn, int n, const double *x, const double *y, double *out)
{
    for (int i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = lm_exp(x[i]); break;
        case 1: out[i] = lm_log(x[i]); break;
        case 2: out[i] = lm_log10(x[i]); break;
        case 3: out[i] = lm_pow(x[i], y[i]); break;
        case 4: out[i] = lm_tanh(x[i]); break;
        case 5: out[i] = lm_expm1(x[i]); break;
        default: out[i] = 0; break;
        }
    }
}
