/*
 * lem_oracle.h — CPU oracle for the PARALLEM per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial (optionally OpenMP) C
 * restatement of the reference's FD -> FA -> MOD timestep, used by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only as
 * the checker.  The product path (parallem_amd/, liblemhip.so) never links,
 * loads or calls anything in this directory.
 *
 * Layout follows the reference: row-major N = W*H cells, Slopes as 8 doubles
 * per cell (Slopes[8*i + d], d = 0..7 = N, NE, E, SE, S, SW, W, NW).
 */
#ifndef LEM_ORACLE_H
#define LEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int W, H;
    double cellsize;
    int flowtype;               /* 0 = SFD, 1 = MFD (default.par:11) */
    double stoneval, fineval, soilTval, soilMval, soilBval, nutval;
    int soil_struct, profile_permeability;
    double Qthreshold, Gslope, kappa, Gflowlim, Flow_A, Flow_B, max_incision;
    double ddk, dck, dgk, temp_inflex, temp_sens;
    double rain, temp;          /* scalar climate (runin rain/temp) */
    int runiniter, start_iter;  /* only used for the FA first-iteration outlet rule */
} orc_params;

typedef struct {
    /* persistent state (N unless noted) */
    double *dem, *Slopes /* 8N */, *SlopePtr;
    double *soilT, *soilM, *stone, *fines, *nut, *soilB, *TotB;
    double *rain, *temp, *last_rain, *last_temp;
    int *mask;
    /* per-step products */
    int *fd, *SFD;
    double *w, *fa, *ero, *inci, *geli, *depo, *dz, *wC, *wP;
    int outlet;
    /* per-step scalars */
    double FA_max, totE, totI, totG, totD, totwC, totwP, totbio;
    int sinkcount;        /* as written to the budgets CSV */
    int sinks_pass1;      /* SFD==0 mask cells after pass-1 routing */
    int plateau_rounds1, plateau_rounds2, nlabels, nedges;
    int fa_uncomputed, sed_uncomputed;
    double totwater;      /* calcwater total of the last runoff pass */
} orc_state;

/* ---- initialisation (M2) ---- */
void orc_createmask(int N, const double *dem, int *mask);
int  orc_init_setoutletcell(int W, int H, const int *mask, const double *dem);
void orc_createSoilTfromformula(int N, const double *dem, double *soilT);
void orc_SetClimateMatrices(const orc_params *p, const int *mask, const double *dem,
                            double rain_scalar, double temp_scalar, double *rainmat, double *tempmat);
void orc_setProcessMatrices(const orc_params *p, orc_state *s);

/* ---- flow directions and routing ---- */
void orc_flowdirs(int W, int H, int cell_size, const double *dem, const int *mask,
                  double *Slopes, int *SFD, int *mfd, int *sinkcounter, int *flatcounter);
int  orc_route_plateaus(int W, int H, const double *dem, const int *mask,
                        double *Slopes, int *fd, int *SFD, float *sp);
int  orc_route_plateaus_literal(int W, int H, const double *dem, const int *mask,
                                double *Slopes, int *fd, int *SFD, float *sp);
int  orc_watershed_labels(int W, int H, const double *dem, const int *mask, const int *SFD, int *ws);
int  orc_spill_heights(int W, int H, const double *dem, const int *mask, const int *ws,
                       int nlab, double *raise, int *sinkcount, int *nedges);
void orc_flow_direction(const orc_params *p, orc_state *s, int iter);

/* ---- runoff and flow accumulation ---- */
void orc_runoff(const orc_params *p, orc_state *s);
void orc_flow_accum(const orc_params *p, orc_state *s, int iter);
int  orc_flow_accum_literal(int W, int H, const int *mask, const int *fd, const double *w, double *fa);

/* ---- MOD (C1-C11) ---- */
void orc_erosion(const orc_params *p, orc_state *s, int iter);
int  orc_sediment_literal(const orc_params *p, orc_state *s);

/* ---- whole step (lem.cpp:83-97) ---- */
void orc_step(const orc_params *p, orc_state *s, int iter);

/* ---- math passthrough for tests ---- */
void orc_math_eval(int fn, int n, const double *x, const double *y, double *out);

/* ---- OpenMP thread control for the CPU baseline ---- */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
