// lem_main.cpp — the reference's driver program (lem.cpp:14-128) on Linux,
// over include/lemhip_compat.h: read the par file, the climate series and
// the DEM, set up the host matrices, make the device context, then run the
// per-step phases in the reference's order with the budgets CSV and the
// periodic grids written as it writes them.
//
//   lem_driver [parfile] [--steps N] [--fp32] [--soilT-nonneg]
//              [--checkpoint-every K] [--restart ITER] [--dump DIR]
//
// The reference ignores its par file's start/final/run-in counts and always
// runs iterations -1000 .. -1 (lem.cpp:64-71); --steps N runs the first N of
// them.  --fp32 selects the FP32 + compensated form (the default is the
// reference's FP64 arithmetic).  --soilT-nonneg clamps the formula soil
// thickness at 0.  --checkpoint-every K saves the device state every K
// iterations under <modelcode>/restart/; --restart ITER resumes from the
// state saved after iteration ITER-1 and continues at ITER.  --dump DIR writes
// the final fields as raw little-endian arrays (dem, fa, soilT, stone .f64;
// fd .i32) for comparison by tests.
#include <errno.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <string>

#include "../../include/lemhip_compat.h"

static int dump(Data *data, const char *dir)
{
    mkdir(dir, 0755);
    int rc = sync_host_arrays(data);
    if (rc != 0) return rc;
    const long n = (long)data->mapInfo.width * data->mapInfo.height;
    const struct { const char *name; const void *a; size_t es; } out[] = {
        {"dem.f64", data->dem, 8}, {"fa.f64", data->fa, 8}, {"soilT.f64", data->soilTPtr, 8},
        {"stone.f64", data->stonePtr, 8}, {"fd.i32", data->fd, 4}};
    for (auto &o : out) {
        std::string path = std::string(dir) + "/" + o.name;
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) return -errno;
        size_t w = fwrite(o.a, o.es, n, f);
        fclose(f);
        if (w != (size_t)n) return -EIO;
    }
    return 0;
}

int main(int argc, char *argv[])
{
    const char *pars = "default.par";
    int steps = 1000, checkpoint = 0, restart_at = 0, have_restart = 0, nonneg = 0, fp32 = 0;
    const char *dumpdir = nullptr;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--steps") && a + 1 < argc) steps = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--fp32")) fp32 = 1;
        else if (!strcmp(argv[a], "--soilT-nonneg")) nonneg = 1;
        else if (!strcmp(argv[a], "--checkpoint-every") && a + 1 < argc) checkpoint = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--restart") && a + 1 < argc) { restart_at = atoi(argv[++a]); have_restart = 1; }
        else if (!strcmp(argv[a], "--dump") && a + 1 < argc) dumpdir = argv[++a];
        else if (argv[a][0] != '-') pars = argv[a];
        else { fprintf(stderr, "unknown option %s\n", argv[a]); return 2; }
    }

    Data data, device;
    memset(&data, 0, sizeof data);
    memset(&device, 0, sizeof device);
    data.precision = fp32 ? LEMHIP_FP32_KAHAN : LEMHIP_FP64;
    createfilenamespace(&data);
    if (readinpar(&data, pars) < 0) return 1;
    checkparread(&data);
    mkdir(data.matrixDIR, 0755);
    data.outlog = fopen(data.logfileOut, "w");
    if (!data.outlog) {
        fprintf(stderr, "cannot open file '%s': %s\n", data.logfileOut, strerror(errno));
        return 1;
    }
    fprintf(data.outlog, "Landscape file   =  %s\nClimate file     =  %s \nBedrock file   =  %s\n", data.demfile,
            data.clim_file, data.bedrockfile);
    double mfree = 0, mtotal = 0;
    if (lemhip_mem_info(0, &mfree, &mtotal) == 0)
        fprintf(data.outlog, "Memory on the device free at start: %.0f KiB total: %.0f KiB\n", mfree / 1024,
                mtotal / 1024);
    const clock_t t0 = clock();
    if (ReadClimateDataFromFile(&data) != 0) fprintf(data.outlog, "no climate series read\n");
    if (readgrid(&data) != 0) return 1;
    createmask(&data);
    init_setoutletcell(&data);
    createSoilTfromformula(&data);
    if (nonneg)
        for (long i = 0; i < (long)data.mapInfo.width * data.mapInfo.height; i++)
            if (data.soilTPtr[i] < 0) data.soilTPtr[i] = 0;
    copyMask(&data, &device);
    SetClimateMatrices(&data, -10000);
    copylastclimate(&data, &device);
    if (createProcessMatrices(&data) != 0) return 1;
    setProcessMatrices(&data);
    if (createDeviceSpace(&data, &device) != 0) return 1;

    const int first = -1000;
    data.start_iter = first;
    int i0 = first;
    if (have_restart) {
        if (retriveProcessMatrices(&data, restart_at - 1) != 0) return 1;
        i0 = restart_at;
    }
    const int last = first + steps < 0 ? first + steps : 0;
    int status = 0;
    for (int i = i0; i < last && status == 0; i++) {
        fprintf(data.outlog, "\n***********************\nStarting mode %s Iteration %d\n***********************\n",
                data.modelcode, i);
        if (i >= 0) SetClimateMatrices(&data, i);
        status |= setdevicespace_FD(&data, &device);
        status |= cuFlowDirection(&data, &device, i);
        status |= cleardevicespace_FD(&data, &device);
        status |= setdevicespace_FA(&data, &device);
        status |= computeRunOffWeights(&data, &device);
        calcwater(&data);
        status |= calcpropsSFD(&data, &device);
        status |= correctmfdflow(&data, &device, i);
        status |= cleardevicespace_FA(&data, &device);
        status |= setdevicespace_Process(&data, &device);
        status |= erosionGPU(&data, &device, i);
        status |= cleardevicespace_Process(&data, &device);
        status |= writeSummaryDataToFile(&data, i);
        zerogrids(&data);
        fprintf(data.outlog, "Finished Iteration %d\n", i);
        fflush(data.outlog);
        status |= writegrids(&data, i);
        if (checkpoint > 0 && (i - first + 1) % checkpoint == 0) status |= savegrids_restart(&data, i);
    }
    if (status == 0 && dumpdir) status = dump(&data, dumpdir);
    const double secs = (double)(clock() - t0) / CLOCKS_PER_SEC;
    fprintf(data.outlog, "Total simulation time of : %20.17lf \n\n", secs);
    printf("Total simulation time of : %20.17lf seconds \n\n", secs);
    clearDeviceSpace(&data, &device);
    deleteProcessMatrices(&data);
    free(data.dem);
    fclose(data.outlog);
    return status == 0 ? 0 : 1;
}
