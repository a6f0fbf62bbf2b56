// lem_compat.cpp — the reference driver's phase and set-up functions
// (include/lemhip_compat.h) over the lemhip C-ABI.  Host code only: the
// device work is liblemhip.so's.
#include <errno.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <string>
#include <vector>

#include "../../include/lemhip_compat.h"

static long ncells(const Data *d) { return (long)d->mapInfo.width * d->mapInfo.height; }

static int fail(Data *data, const char *what, int rc)
{
    const char *msg = data->ctx ? lemhip_last_error(data->ctx) : "";
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, msg);
    if (data->outlog) fprintf(data->outlog, "%s failed (%d): %s\n", what, rc, msg);
    return rc;
}

#define LEM_CALL(data, call)                       \
    do {                                           \
        int rc_ = (call);                          \
        if (rc_ != 0) return fail(data, #call, rc_); \
    } while (0)

// ------------------------------------------------------------------ set-up

void createfilenamespace(Data *data)
{
    // the reference mallocs its file-name buffers here (memory.cu); the
    // names live inside Data here
    (void)data;
}

// createProcessMatrices (memory.cu:89-128): host arrays of the grid
int createProcessMatrices(Data *data)
{
    const long n = ncells(data);
    double **dfields[] = {&data->fa, &data->runoffweight, &data->SlopePtr, &data->finesPtr, &data->stonePtr,
                          &data->soilTPtr, &data->soilMPtr, &data->soilBPtr, &data->nutPtr, &data->TotBPtr,
                          &data->eroPtr, &data->inciPtr, &data->geliPtr, &data->depoPtr, &data->weatherC,
                          &data->weatherP, &data->dz, &data->rainmat, &data->tempmat, &data->last_rainmat,
                          &data->last_tempmat};
    for (double **f : dfields) {
        if (*f) continue;
        *f = (double *)calloc(n, sizeof(double));
        if (!*f) return -1;
    }
    if (!data->Slopes && !(data->Slopes = (double *)calloc(8 * n, sizeof(double)))) return -1;
    int **ifields[] = {&data->mask, &data->fd, &data->SFD};
    for (int **f : ifields) {
        if (*f) continue;
        *f = (int *)calloc(n, sizeof(int));
        if (!*f) return -1;
    }
    return 0;
}

void deleteProcessMatrices(Data *data)
{
    void *ptrs[] = {data->fa, data->runoffweight, data->SlopePtr, data->finesPtr, data->stonePtr, data->soilTPtr,
                    data->soilMPtr, data->soilBPtr, data->nutPtr, data->TotBPtr, data->eroPtr, data->inciPtr,
                    data->geliPtr, data->depoPtr, data->weatherC, data->weatherP, data->dz, data->rainmat,
                    data->tempmat, data->last_rainmat, data->last_tempmat, data->Slopes, data->mask, data->fd,
                    data->SFD, data->rainchange, data->tempchange, data->year};
    for (void *p : ptrs) free(p);
}

// createmask (memory.cu:217-242)
void createmask(Data *data)
{
    if (!data->mask) data->mask = (int *)calloc(ncells(data), sizeof(int));
    for (long i = 0; i < ncells(data); i++) data->mask[i] = data->dem[i] == -9999 ? 0 : 1;
}

// init_setoutletcell (Data.cpp:38-66): the first cell holding the strict
// minimum of the mask cells, starting from 1000
void init_setoutletcell(Data *data)
{
    double lowest = 1000.;
    int index = 0;
    for (long i = 0; i < ncells(data); i++)
        if (data->mask[i] == 1 && data->dem[i] < lowest) {
            lowest = data->dem[i];
            index = (int)i;
        }
    data->outletcellidx = index;
    if (data->outlog)
        fprintf(data->outlog, "identified outlet cell has index %d and value %f \n", index, data->dem[index]);
}

// createSoilTfromformula (memory.cu:197-214)
void createSoilTfromformula(Data *data)
{
    if (!data->soilTPtr) data->soilTPtr = (double *)calloc(ncells(data), sizeof(double));
    for (long i = 0; i < ncells(data); i++)
        data->soilTPtr[i] = data->dem[i] > 0 ? ((data->dem[i] - 400) / 1400) * 5 : 0.0;
}

// SetClimateMatrices (Data.cpp:211-303): the previous scalars become the
// last climate; from iteration 0 the series' values replace rain and temp.
// Rain doubles down the rows, temperature follows a lapse rate of 6.5 K/km
// from 467.228 m.  The set-up call (-10000) also fills the last-climate
// matrices; later calls send the new climate to the device.
static void fill_climate(Data *data, double rain, double temp, double *rmat, double *tmat)
{
    const int W = data->mapInfo.width, H = data->mapInfo.height;
    const double increment = (rain * 2 - rain) / H;
    for (int r = 0; r < H; r++) {
        const double therain = rain + r * increment;
        for (int c = 0; c < W; c++) {
            const long i = (long)r * W + c;
            if (data->mask[i] == 1) {
                rmat[i] = therain;
                tmat[i] = temp - ((data->dem[i] - 467.228) * 0.0065);
            } else {
                rmat[i] = -9999.;
                tmat[i] = -9999.;
            }
        }
    }
}

void SetClimateMatrices(Data *data, int iteration)
{
    const long n = ncells(data);
    double **mats[] = {&data->rainmat, &data->tempmat, &data->last_rainmat, &data->last_tempmat};
    for (double **m : mats)
        if (!*m) *m = (double *)calloc(n, sizeof(double));
    data->last_rain = data->rain;
    data->last_temp = data->temp;
    if (iteration > -10000 && iteration >= 0 && data->rainchange && iteration < data->max_iterations) {
        data->rain = data->rainchange[iteration];
        data->temp = data->tempchange[iteration];
    }
    fill_climate(data, data->rain, data->temp, data->rainmat, data->tempmat);
    if (iteration == -10000) fill_climate(data, data->last_rain, data->last_temp, data->last_rainmat, data->last_tempmat);
    else if (data->ctx) (void)lemhip_set_climate(data->ctx, data->rain, data->temp);
}

// setProcessMatrices (Data.cpp:115-185), values from the par file
void setProcessMatrices(Data *data)
{
    const double rescale = 10000 / (data->mapInfo.cellsize * data->mapInfo.cellsize);
    for (long x = 0; x < ncells(data); x++) {
        const bool m = data->mask[x] == 1;
        data->fd[x] = 0;
        data->SFD[x] = 0;
        data->fa[x] = 0;
        data->SlopePtr[x] = 0;
        data->dz[x] = 0;
        data->runoffweight[x] = m ? 1.0 : 0.0;
        data->stonePtr[x] = m ? data->stoneval : 0;
        data->finesPtr[x] = m ? data->fineval : 0;
        data->soilMPtr[x] = m ? data->soilMval : 0;
        data->nutPtr[x] = m ? data->nutval : 0;
        data->soilBPtr[x] = m ? data->soilBval : 0;
        if (!m) data->soilTPtr[x] = 0;
        data->TotBPtr[x] = m ? (3000.0 + 5.0 * (data->rainmat[x] - (100 + 100 / (data->tempmat[x] + 15)))) / rescale
                             : 0;
        data->eroPtr[x] = data->inciPtr[x] = data->geliPtr[x] = data->depoPtr[x] = 0;
        data->weatherC[x] = data->weatherP[x] = 0;
    }
    memset(data->Slopes, 0, sizeof(double) * 8 * ncells(data));
}

// copyMask / copylastclimate (memory_dev.cu:242-289): the mask and the
// climate fields go to the device with everything else in createDeviceSpace
int copyMask(Data *data, Data *device)
{
    (void)data;
    (void)device;
    return 0;
}

int copylastclimate(Data *data, Data *device)
{
    (void)data;
    (void)device;
    return 0;
}

// createDeviceSpace (memory_dev.cu:293-321): one context holding every field
// for the whole run, initialised from the host arrays
int createDeviceSpace(Data *data, Data *device)
{
    lem_params p;
    memset(&p, 0, sizeof p);
    p.width = data->mapInfo.width;
    p.height = data->mapInfo.height;
    p.cellsize = data->mapInfo.cellsize;
    p.flowdirectiontype = data->flowdirectiontype;
    p.stoneval = data->stoneval;
    p.fineval = data->fineval;
    p.soilTval = data->soilTval;
    p.soilMval = data->soilMval;
    p.soilBval = data->soilBval;
    p.nutval = data->nutval;
    p.soil_struct = data->soil_struct;
    p.profile_permeability = data->profile_permeability;
    p.Qthreshold = data->Qthreshold;
    p.Gslope = data->Gslope;
    p.kappa = data->kappa;
    p.Gflowlim = data->Gflowlim;
    p.Flow_A = data->Flow_A;
    p.Flow_B = data->Flow_B;
    p.max_incision = data->max_incision;
    p.ddk = data->ddk;
    p.dck = data->dck;
    p.dgk = data->dgk;
    p.temp_inflex = data->temp_inflex;
    p.temp_sens = data->temp_sens;
    p.rain = data->rain;
    p.temp = data->temp;
    p.runiniter = data->runiniter;
    // the first-iteration outlet rule (mfd_simple.cu:336) compares with
    // start_iter - runiniter at the time of the step, and the driver has set
    // start_iter to -1000 by then (lem.cpp:68)
    p.start_iter = -1000;
    lemhip_ctx *ctx = nullptr;
    int rc = lemhip_create(&p, data->precision, 0, &ctx);
    if (rc != 0) {
        fprintf(stderr, "lemhip_create failed (%d): %s\n", rc, ctx ? lemhip_last_error(ctx) : "");
        if (ctx) lemhip_destroy(ctx);
        return rc;
    }
    data->ctx = device->ctx = ctx;
    device->precision = data->precision;
    device->mapInfo = data->mapInfo;
    // the device derives the same mask, outlet, formula soil thickness and
    // process matrices from the surface and rain field; the host's soil
    // thickness, climate and process matrices are then uploaded over them, so
    // whatever the caller set before this call (a restart, a bedrock file)
    // is what the run starts from
    LEM_CALL(data, lemhip_init_state(ctx, data->dem, data->rainmat, 0));
    LEM_CALL(data, lemhip_upload(ctx, LEM_F_SOILT, data->soilTPtr));
    const struct { int f; const double *a; } up[] = {
        {LEM_F_STONE, data->stonePtr}, {LEM_F_FINES, data->finesPtr}, {LEM_F_SOILM, data->soilMPtr},
        {LEM_F_SOILB, data->soilBPtr}, {LEM_F_NUT, data->nutPtr}, {LEM_F_TOTB, data->TotBPtr},
        {LEM_F_TEMP, data->tempmat}, {LEM_F_LAST_RAIN, data->last_rainmat}, {LEM_F_LAST_TEMP, data->last_tempmat},
        {LEM_F_SLOPES, data->Slopes}};
    for (auto &u : up) LEM_CALL(data, lemhip_upload(ctx, u.f, u.a));
    LEM_CALL(data, lemhip_set_outlet(ctx, data->outletcellidx));
    return 0;
}

int clearDeviceSpace(Data *data, Data *device)
{
    int rc = data->ctx ? lemhip_destroy(data->ctx) : 0;
    data->ctx = device->ctx = nullptr;
    return rc;
}

// ----------------------------------------------------------------- phases

// every field stays resident: the per-phase allocate/copy/free calls of the
// reference (memory_dev.cu:33-240) have nothing to do
int setdevicespace_FD(Data *, Data *) { return 0; }
int cleardevicespace_FD(Data *, Data *) { return 0; }
int setdevicespace_FA(Data *, Data *) { return 0; }
int cleardevicespace_FA(Data *, Data *) { return 0; }
int setdevicespace_Process(Data *, Data *) { return 0; }
int cleardevicespace_Process(Data *, Data *) { return 0; }
void zerogrids(Data *) {}

// prop = 1 for every mask cell (runoffweight.cu:115-138) is folded into the
// accumulation pass
int calcpropsSFD(Data *, Data *) { return 0; }

static int read_totals(Data *data)
{
    lem_summary s;
    LEM_CALL(data, lemhip_summary(data->ctx, &s));
    data->FA_max = s.FA_max;
    data->totE = s.totE;
    data->totI = s.totI;
    data->totG = s.totG;
    data->totD = s.totD;
    data->totweatherC = s.totweatherC;
    data->totweatherP = s.totweatherP;
    data->totbio = s.totbio;
    data->totwater = s.totwater;
    data->sinkcount = s.sinkcount;
    data->outletcellidx = s.outletcellidx;
    return 0;
}

int cuFlowDirection(Data *data, Data *device, int iter)
{
    LEM_CALL(data, lemhip_flow_direction(device->ctx, iter));
    return 0;
}

int computeRunOffWeights(Data *data, Data *device)
{
    LEM_CALL(data, lemhip_runoff(device->ctx));
    return 0;
}

// calcwater (runoffweight.cu:356-377): the total runoff weight, logged
void calcwater(Data *data)
{
    lem_summary s;
    if (lemhip_summary(data->ctx, &s) != 0) return;
    data->totwater = s.totwater;
    printf("Total amount of water in grid for runoff = %lf \n", s.totwater);
    if (data->outlog) fprintf(data->outlog, "Total amount of water in grid for runoff = %lf \n", s.totwater);
}

int correctmfdflow(Data *data, Data *device, int iter)
{
    LEM_CALL(data, lemhip_flow_accum(device->ctx, iter));
    return read_totals(data);
}

int erosionGPU(Data *data, Data *device, int iter)
{
    LEM_CALL(data, lemhip_erosion(device->ctx, iter));
    return read_totals(data);
}

int copytolastclimate(Data *data, Data *device)
{
    LEM_CALL(data, lemhip_copy_to_last_climate(device->ctx));
    return 0;
}

int sync_host_arrays(Data *data)
{
    lemhip_ctx *c = data->ctx;
    const struct { int f; void *a; } down[] = {
        {LEM_F_DEM, data->dem},         {LEM_F_MASK, data->mask},       {LEM_F_FD, data->fd},
        {LEM_F_SFD, data->SFD},         {LEM_F_FA, data->fa},           {LEM_F_RUNOFF, data->runoffweight},
        {LEM_F_SLOPEPTR, data->SlopePtr}, {LEM_F_SLOPES, data->Slopes}, {LEM_F_FINES, data->finesPtr},
        {LEM_F_STONE, data->stonePtr},  {LEM_F_SOILT, data->soilTPtr},  {LEM_F_SOILM, data->soilMPtr},
        {LEM_F_SOILB, data->soilBPtr},  {LEM_F_NUT, data->nutPtr},      {LEM_F_TOTB, data->TotBPtr},
        {LEM_F_ERO, data->eroPtr},      {LEM_F_INCI, data->inciPtr},    {LEM_F_GELI, data->geliPtr},
        {LEM_F_DEPO, data->depoPtr},    {LEM_F_WC, data->weatherC},     {LEM_F_WP, data->weatherP},
        {LEM_F_DZ, data->dz},           {LEM_F_RAIN, data->rainmat},    {LEM_F_TEMP, data->tempmat}};
    for (auto &d : down) LEM_CALL(data, lemhip_download(c, d.f, d.a));
    return 0;
}

// writegrids (io.cpp:201-230): fd, fa and height every 50 iterations,
// erosion, incision, gelifluction and deposition every 100, as
// <matrixDIR>/<iteration>_<name>.asc, after the device state is copied to
// the host arrays (the reference's phases copy back after every kernel)
int writegrids(Data *data, int iteration)
{
    const bool k50 = iteration % 50 == 0, k100 = iteration % 100 == 0;
    if (!k50 && !k100) return 0;
    int rc = sync_host_arrays(data);
    if (rc != 0) return rc;
    auto name = [&](const char *what) {
        return std::string(data->matrixDIR) + "/" + std::to_string(iteration) + "_" + what + ".asc";
    };
    if (k50) {
        if ((rc = write_int(data, data->fd, name("FD").c_str())) != 0) return rc;
        if ((rc = write_double(data, data->fa, name("FA").c_str())) != 0) return rc;
        if ((rc = write_double(data, data->dem, name("height").c_str())) != 0) return rc;
    }
    if (k100) {
        if ((rc = write_double(data, data->eroPtr, name("ero").c_str())) != 0) return rc;
        if ((rc = write_double(data, data->inciPtr, name("inci").c_str())) != 0) return rc;
        if ((rc = write_double(data, data->geliPtr, name("geli").c_str())) != 0) return rc;
        if ((rc = write_double(data, data->depoPtr, name("depo").c_str())) != 0) return rc;
    }
    return 0;
}

// ------------------------------------------------------ checkpoint / restart

static std::string snap_path(const Data *data, int iteration)
{
    return std::string(data->matrixDIR) + "/restart/" + std::to_string(iteration) + "_state.lemsnap";
}

int savegrids_restart(Data *data, int iteration)
{
    size_t n = 0;
    LEM_CALL(data, lemhip_snapshot_size(data->ctx, 0, &n));
    std::vector<char> buf(n);
    LEM_CALL(data, lemhip_snapshot(data->ctx, 0, buf.data(), n));
    std::string dir = std::string(data->matrixDIR) + "/restart";
    mkdir(dir.c_str(), 0755);
    const std::string path = snap_path(data, iteration);
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return fail(data, path.c_str(), -errno);
    const bool ok = fwrite(buf.data(), 1, n, f) == n;
    if (fclose(f) != 0 || !ok) return fail(data, path.c_str(), -EIO);
    return 0;
}

int retriveProcessMatrices(Data *data, int iteration)
{
    const std::string path = snap_path(data, iteration);
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return fail(data, path.c_str(), -errno);
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    fclose(f);
    LEM_CALL(data, lemhip_restore(data->ctx, 0, buf.data(), buf.size()));
    int o = 0;
    LEM_CALL(data, lemhip_get_outlet(data->ctx, &o));
    data->outletcellidx = o;
    return sync_host_arrays(data);
}
