/*
 * lemhip_compat.h — the reference driver's C++ interface over liblemhip.
 *
 * PARALLEM's driver (lem.cpp:14-128) passes two `Data` structs to free
 * functions: `data` with host arrays and `device` with raw device pointers
 * (Data.h:15-187, memory_dev.h:37-57).  This header keeps those names and
 * call shapes so a lem.cpp-shaped main loop compiles against it unchanged:
 *
 *   - `Data` keeps the host fields the driver and its I/O touch, with the
 *     reference's names (dem, fd, SFD, fa, runoffweight, SlopePtr, Slopes,
 *     finesPtr, stonePtr, soilTPtr, soilMPtr, soilBPtr, nutPtr, TotBPtr,
 *     eroPtr, inciPtr, geliPtr, depoPtr, weatherC, weatherP, dz, rainmat,
 *     tempmat, last_rainmat, last_tempmat, mask, mapInfo, the par values and
 *     the per-step totals); its `ctx` member stands in for the device mirror:
 *     createDeviceSpace makes the lemhip context that holds every field in
 *     HBM and stores it in both structs, so functions that take only `data`
 *     (writegrids, the restart pair) reach it too.
 *   - setdevicespace_* / cleardevicespace_* / zerogrids become no-ops: state
 *     stays resident across phases and steps.
 *   - the phase functions forward to the C-ABI and copy back the per-step
 *     totals the budgets CSV reads; host arrays are refreshed by writegrids
 *     (and by sync_host_arrays for callers that read them directly).
 *
 * Host-side initialisation (createmask, init_setoutletcell,
 * createSoilTfromformula, SetClimateMatrices, setProcessMatrices) is kept
 * with the reference's semantics so the host arrays the driver owns hold the
 * same values as the device after createDeviceSpace.
 *
 * The I/O of the reference driver (io.cpp) is declared here too: ESRI ASCII
 * grids (readgrid, write_double, write_int, write_float), writegrids, the
 * 35-line parameter file (readinpar), the climate file
 * (ReadClimateDataFromFile) and the budgets CSV (writeSummaryDataToFile).
 * Errors are returned, never turned into exit(): the reference's readgrid
 * and readinpar call exit(0) on a missing file, these return non-zero.
 */
#ifndef LEMHIP_COMPAT_H
#define LEMHIP_COMPAT_H

#include <stdio.h>
#include "lemhip.h"

struct MapInfo {              /* MapInfo.h:4-12 */
    int width, height;
    double xllcorner, yllcorner, cellsize, nodata;
};

struct Data {
    /* file names and run control (readinpar, io.cpp:376-580) */
    char modelcode[64];
    char demfile[256], clim_file[256], bedrockfile[256];
    char matrixDIR[256], logfileOut[512], outfilename[512];
    FILE *outlog;
    int restart, start_iter, max_iterations, runiniter, flowdirectiontype;
    /* par values (default.par:14-35) */
    double rain, temp, last_rain, last_temp;
    double stoneval, fineval, soilTval, soilMval, soilBval, nutval;
    int soil_struct, profile_permeability;
    double Qthreshold, Gslope, kappa, Gflowlim, Flow_A, Flow_B, max_incision, dummy3, dummy4;
    double ddk, dck, dgk, temp_inflex, temp_sens;
    /* climate series (ReadClimateDataFromFile, io.cpp:265-319) */
    double *rainchange, *tempchange;
    int *year;
    /* grid */
    MapInfo mapInfo;
    double *dem;
    int *mask, *fd, *SFD;
    double *fa, *runoffweight, *SlopePtr, *Slopes;
    double *finesPtr, *stonePtr, *soilTPtr, *soilMPtr, *soilBPtr, *nutPtr, *TotBPtr;
    double *eroPtr, *inciPtr, *geliPtr, *depoPtr, *weatherC, *weatherP, *dz;
    double *rainmat, *tempmat, *last_rainmat, *last_tempmat;
    /* per-step scalars (written by the phases, read by the CSV writer) */
    int outletcellidx, sinkcount;
    double FA_max, totE, totI, totG, totD, totbio, totweatherC, totweatherP, totwater;
    /* the device mirror: the lemhip context */
    lemhip_ctx *ctx;
    int precision;            /* LEMHIP_FP64 (the reference's arithmetic) or LEMHIP_FP32_KAHAN */
};

/* ---- set-up (memory.cu, Data.cpp, memory_dev.cu) ---- */
void createfilenamespace(Data *data);
int createProcessMatrices(Data *data);
void deleteProcessMatrices(Data *data);
void createmask(Data *data);
void init_setoutletcell(Data *data);
void createSoilTfromformula(Data *data);
void SetClimateMatrices(Data *data, int iteration);
void setProcessMatrices(Data *data);
int copyMask(Data *data, Data *device);
int copylastclimate(Data *data, Data *device);
int createDeviceSpace(Data *data, Data *device);
int clearDeviceSpace(Data *data, Data *device);

/* ---- per-step phases (lem.cpp:83-97) ---- */
int setdevicespace_FD(Data *data, Data *device);
int cleardevicespace_FD(Data *data, Data *device);
int cuFlowDirection(Data *data, Data *device, int iter);
int setdevicespace_FA(Data *data, Data *device);
int cleardevicespace_FA(Data *data, Data *device);
int computeRunOffWeights(Data *data, Data *device);
void calcwater(Data *data);
int calcpropsSFD(Data *data, Data *device);
int correctmfdflow(Data *data, Data *device, int iter);
int setdevicespace_Process(Data *data, Data *device);
int cleardevicespace_Process(Data *data, Data *device);
int erosionGPU(Data *data, Data *device, int iter);
int copytolastclimate(Data *data, Data *device);
void zerogrids(Data *data);
/* copies the device state into the host arrays of `data` (no reference
 * counterpart: there every phase copies its products back) */
int sync_host_arrays(Data *data);

/* ---- checkpoint / restart (retriveProcessMatrices, ioGDAL.cpp:662-804) ----
 * save: the device state to <matrixDIR>/restart/<iteration>_state.lemsnap;
 * retrieve: the same file back into the context (after createDeviceSpace). */
int savegrids_restart(Data *data, int iteration);
int retriveProcessMatrices(Data *data, int iteration);

/* ---- I/O (io.cpp) ---- */
int readgrid(Data *data);
int write_double(Data *data, const double *grid, const char *filename);
int write_int(Data *data, const int *grid, const char *filename);
int write_float(Data *data, const float *grid, const char *filename);
int writegrids(Data *data, int iteration);
int writeSummaryDataToFile(Data *data, int iteration);
int ReadClimateDataFromFile(Data *data);
int readinpar(Data *data, const char *file);
int checkparread(Data *data);

#endif
