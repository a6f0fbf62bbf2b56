// lem_io.cpp — the reference driver's file formats (io.cpp) over the Data
// struct of include/lemhip_compat.h: ESRI ASCII grids, the 35-line
// positional parameter file, the climate series and the budgets CSV.
// Host code only (no HIP); built into the driver and, for the CPU tests,
// into liblemio.so with the lemio_* entry points at the end of this file.
#include <ctype.h>
#include <errno.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <string>
#include <vector>

#include "../../include/lemhip_compat.h"

// ------------------------------------------------------------ ESRI ASCII

// Header keys in order (io.cpp:43-55 reads six "<key> <value>" pairs and
// ignores the key text); values are read as numbers whatever their spelling.
static int read_header(FILE *in, MapInfo *mi)
{
    char key[64];
    double v[6];
    for (int k = 0; k < 6; k++)
        if (fscanf(in, "%63s %lf", key, &v[k]) != 2) return -1;
    mi->width = (int)v[0];
    mi->height = (int)v[1];
    mi->xllcorner = v[2];
    mi->yllcorner = v[3];
    mi->cellsize = v[4];
    mi->nodata = v[5];
    return (mi->width > 0 && mi->height > 0) ? 0 : -1;
}

// readgrid (io.cpp:28-96): header then W*H values, row by row from the top
int readgrid(Data *data)
{
    FILE *in = fopen(data->demfile, "r");
    if (!in) {
        fprintf(stderr, "readgrid: cannot open %s: %s\n", data->demfile, strerror(errno));
        return -1;
    }
    if (read_header(in, &data->mapInfo) != 0) {
        fclose(in);
        fprintf(stderr, "readgrid: bad header in %s\n", data->demfile);
        return -2;
    }
    const long n = (long)data->mapInfo.width * data->mapInfo.height;
    free(data->dem);
    data->dem = (double *)malloc(sizeof(double) * n);
    if (!data->dem) { fclose(in); return -3; }
    for (long i = 0; i < n; i++)
        if (fscanf(in, "%lf", &data->dem[i]) != 1) {
            fclose(in);
            fprintf(stderr, "readgrid: %s ends after %ld of %ld values\n", data->demfile, i, n);
            return -4;
        }
    fclose(in);
    if (data->outlog) {
        fprintf(data->outlog, "ncols %d \nnrows %d \nxllcorner %lf \nyllcorner %lf \ncellsize %lf \nnodata %lf \n",
                data->mapInfo.width, data->mapInfo.height, data->mapInfo.xllcorner, data->mapInfo.yllcorner,
                data->mapInfo.cellsize, data->mapInfo.nodata);
        fprintf(data->outlog, "Finished reading %s...\n", data->demfile);
        fflush(data->outlog);
    }
    return 0;
}

// write_double / write_int / write_float (io.cpp:98-199): the same six-line
// header, then one line per row; cells outside the mask are written as -9999
template <typename F>
static int write_grid(const Data *data, const char *filename, F cell)
{
    FILE *out = fopen(filename, "w");
    if (!out) {
        fprintf(stderr, "write grid: cannot open %s: %s\n", filename, strerror(errno));
        return -1;
    }
    const MapInfo &m = data->mapInfo;
    fprintf(out, "ncols %d\nnrows %d\nxllcorner %lf\nyllcorner %lf\ncellsize %lf\nNODATA_value %lf\n", m.width,
            m.height, m.xllcorner, m.yllcorner, m.cellsize, m.nodata);
    for (int r = 0; r < m.height; r++) {
        for (int c = 0; c < m.width; c++) {
            const long i = (long)r * m.width + c;
            if (data->mask[i] == 1) cell(out, i);
            else fputs("-9999 ", out);
        }
        fputc('\n', out);
    }
    return fclose(out) == 0 ? 0 : -2;
}

int write_double(Data *data, const double *grid, const char *filename)
{
    return write_grid(data, filename, [&](FILE *o, long i) { fprintf(o, "%lf ", grid[i]); });
}

int write_int(Data *data, const int *grid, const char *filename)
{
    return write_grid(data, filename, [&](FILE *o, long i) { fprintf(o, "%d ", grid[i]); });
}

int write_float(Data *data, const float *grid, const char *filename)
{
    return write_grid(data, filename, [&](FILE *o, long i) { fprintf(o, "%f ", grid[i]); });
}

// ------------------------------------------------------------ budgets CSV

// writeSummaryDataToFile (io.cpp:233-263): a header line when the iteration
// is start_iter, then one line per iteration with the reference's columns
// and number formats; sedoutflow = (totE + totI + totG - totD) / 1300000 and
// waterout = FA_max * 1000
int writeSummaryDataToFile(Data *data, int iteration)
{
    if (iteration == data->start_iter) {
        FILE *h = fopen(data->outfilename, "w");
        if (!h) return -1;
        fprintf(h, "%s, %s, %s, %s, %s, %s, %s, %s, %s, %s, %s, %s, %s, %s, %s \n", "iteration", "year", "temp",
                "rain", "diffuse", "concentrated", "gelifluction", "chemicalW", "physicalW", "deposition", "totbio",
                "FA_max", "sedoutflow", "waterout", "sinkcount");
        fclose(h);
    }
    FILE *out = fopen(data->outfilename, "a");
    if (!out) return -1;
    const double sedoutflow = (data->totE + data->totI + data->totG - data->totD) / 1300000;
    const double totalrunoff = data->FA_max * 1000;
    if (data->outlog) fprintf(data->outlog, "Ave erosion rate mm/a : %10.8lf \n", sedoutflow);
    // the reference's concentrated-erosion column is printed with "%10.81lf"
    // (81 decimals, io.cpp:254); kept, since readers of the CSV parse it
    fprintf(out, "%d, %d, %f, %f, %10.8lf, %10.81lf, %10.8lf, %10.8lf, %10.8lf, %10.8lf, %10.8lf, %10.3lf, %10.8lf,  %10.5lf, %d \n",
            iteration, iteration, data->temp, data->rain, data->totE, data->totI, data->totG, data->totweatherC,
            data->totweatherP, data->totD, data->totbio, data->FA_max, sedoutflow, totalrunoff, data->sinkcount);
    return fclose(out) == 0 ? 0 : -2;
}

// -------------------------------------------------------------- climate

// ReadClimateDataFromFile (io.cpp:265-319): three heading words, then
// max_iterations rows of "year rain_change temp_change" (read as float, as
// the reference does)
int ReadClimateDataFromFile(Data *data)
{
    const int n = data->max_iterations > 0 ? data->max_iterations : 0;
    free(data->rainchange);
    free(data->tempchange);
    free(data->year);
    data->rainchange = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    data->tempchange = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    data->year = (int *)calloc(n > 0 ? n : 1, sizeof(int));
    FILE *in = fopen(data->clim_file, "r");
    if (!in) {
        fprintf(stderr, "climate file %s: %s\n", data->clim_file, strerror(errno));
        return 1;
    }
    char a[64], b[64], c[64];
    if (fscanf(in, "%63s %63s %63s", a, b, c) != 3) { fclose(in); return 2; }
    for (int i = 0; i < n; i++) {
        int y;
        float dr, dt;
        if (fscanf(in, "%d %f %f", &y, &dr, &dt) != 3) { fclose(in); return 3; }
        data->year[i] = y;
        data->rainchange[i] = (double)dr;
        data->tempchange[i] = (double)dt;
    }
    fclose(in);
    if (data->outlog) fprintf(data->outlog, "climate file read \n");
    return 0;
}

// ------------------------------------------------------------ parameters

// Only the first 38 characters of a line count (io.cpp:400+ copies 38, or 10
// or 20 for numbers, into a 39-character buffer); blanks are dropped from
// file names (stripblank) and the text ends at the newline (choppy)
static std::string field(const char *line, size_t width)
{
    std::string s(line, strnlen(line, width));
    size_t nl = s.find('\n');
    if (nl != std::string::npos) s.resize(nl);
    return s;
}

static void name_field(const char *line, char *dst, size_t cap)
{
    std::string s = field(line, 38), out;
    for (char ch : s)
        if (ch != ' ') out += ch;
    snprintf(dst, cap, "%s", out.c_str());
}

// readinpar (io.cpp:376-580): 35 positional lines — title, date, model code
// (also the output directory), DEM file, climate file, bedrock file,
// restart, start, final and run-in iterations, flow routing type, run-in
// rain and temperature, stone, fines, structure, permeability, soil
// thickness, moisture, organics and nutrients, discharge threshold,
// gelifluction slope threshold, kappa and flow limit, incision A, B and
// maximum, two unused values, the three travel distances, and the two
// vegetation growth parameters
int readinpar(Data *data, const char *file)
{
    FILE *pf = fopen(file, "r");
    if (!pf) {
        fprintf(stderr, "parameter file %s does not exist\n", file);
        return -1;
    }
    std::vector<std::string> lines;
    char buf[256];
    while (lines.size() < 39 && fgets(buf, sizeof buf, pf)) lines.push_back(buf);
    fclose(pf);
    if (lines.size() < 35) {
        fprintf(stderr, "parameter file %s has %zu lines, 35 needed\n", file, lines.size());
        return -2;
    }
    int k = 2;   // title and date
    name_field(lines[k++].c_str(), data->modelcode, sizeof data->modelcode);
    snprintf(data->matrixDIR, sizeof data->matrixDIR, "%s", data->modelcode);
    snprintf(data->logfileOut, sizeof data->logfileOut, "%s/%s_logfile.txt", data->matrixDIR, data->modelcode);
    snprintf(data->outfilename, sizeof data->outfilename, "%s/%s_budgets.csv", data->matrixDIR, data->modelcode);
    name_field(lines[k++].c_str(), data->demfile, sizeof data->demfile);
    name_field(lines[k++].c_str(), data->clim_file, sizeof data->clim_file);
    name_field(lines[k++].c_str(), data->bedrockfile, sizeof data->bedrockfile);
    auto I = [&](size_t w) { return (int)strtol(field(lines[k++].c_str(), w).c_str(), nullptr, 10); };
    auto D = [&]() { return strtod(field(lines[k++].c_str(), 20).c_str(), nullptr); };
    data->restart = I(10);
    data->start_iter = I(10);
    data->max_iterations = I(10);
    data->runiniter = I(10);
    data->flowdirectiontype = I(10);
    data->rain = D();
    data->temp = D();
    data->stoneval = D();
    data->fineval = D();
    data->soil_struct = I(20);
    data->profile_permeability = I(20);
    data->soilTval = D();
    data->soilMval = D();
    data->soilBval = D();
    data->nutval = D();
    data->Qthreshold = D();
    data->Gslope = D();
    data->kappa = D();
    data->Gflowlim = D();
    data->Flow_A = D();
    data->Flow_B = D();
    data->max_incision = D();
    data->dummy3 = D();
    data->dummy4 = D();
    data->ddk = D();
    data->dck = D();
    data->dgk = D();
    data->temp_inflex = D();
    data->temp_sens = D();
    data->last_rain = data->rain;
    data->last_temp = data->temp;
    return k;   // lines used (35)
}

int checkparread(Data *data)
{
    printf("This is model : %s\nLandscape file is: %s\nClimate file is: %s\nLog file is: %s\n"
           "Summary csv file is: %s\n\n",
           data->modelcode, data->demfile, data->clim_file, data->logfileOut, data->outfilename);
    printf("restart %d start %d finish %d runin %d flow type %d rain %f temp %f\n", data->restart, data->start_iter,
           data->max_iterations, data->runiniter, data->flowdirectiontype, data->rain, data->temp);
    return 1;
}

// ------------------------------------------------------ test entry points

extern "C" {

// parameter file -> the numbers in file order (35 - 6 text lines = 29
// values; `out` holds 29 doubles) and the four names; returns lines used
int lemio_read_par(const char *file, double *out, char *modelcode, char *demfile, char *climfile,
                   char *bedrockfile)
{
    Data d;
    memset(&d, 0, sizeof d);
    int rc = readinpar(&d, file);
    if (rc < 0) return rc;
    const double v[29] = {(double)d.restart, (double)d.start_iter, (double)d.max_iterations, (double)d.runiniter,
                          (double)d.flowdirectiontype, d.rain, d.temp, d.stoneval, d.fineval,
                          (double)d.soil_struct, (double)d.profile_permeability, d.soilTval, d.soilMval,
                          d.soilBval, d.nutval, d.Qthreshold, d.Gslope, d.kappa, d.Gflowlim, d.Flow_A, d.Flow_B,
                          d.max_incision, d.dummy3, d.dummy4, d.ddk, d.dck, d.dgk, d.temp_inflex, d.temp_sens};
    memcpy(out, v, sizeof v);
    strcpy(modelcode, d.modelcode);
    strcpy(demfile, d.demfile);
    strcpy(climfile, d.clim_file);
    strcpy(bedrockfile, d.bedrockfile);
    return rc;
}

// ESRI ASCII grid -> header (w, h, xll, yll, cs, nodata) and values
int lemio_read_grid(const char *file, int *wh, double *hdr4, double *values, long cap)
{
    Data d;
    memset(&d, 0, sizeof d);
    snprintf(d.demfile, sizeof d.demfile, "%s", file);
    int rc = readgrid(&d);
    if (rc != 0) return rc;
    wh[0] = d.mapInfo.width;
    wh[1] = d.mapInfo.height;
    hdr4[0] = d.mapInfo.xllcorner;
    hdr4[1] = d.mapInfo.yllcorner;
    hdr4[2] = d.mapInfo.cellsize;
    hdr4[3] = d.mapInfo.nodata;
    const long n = (long)wh[0] * wh[1];
    if (values && cap >= n) memcpy(values, d.dem, n * sizeof(double));
    free(d.dem);
    return 0;
}

// grid -> ESRI ASCII (kind 0 double "%lf", 1 int, 2 float), masked cells -9999
int lemio_write_grid(const char *file, int kind, int w, int h, const double *hdr4, const int *mask,
                     const void *values)
{
    Data d;
    memset(&d, 0, sizeof d);
    d.mapInfo = MapInfo{w, h, hdr4[0], hdr4[1], hdr4[2], hdr4[3]};
    d.mask = const_cast<int *>(mask);
    if (kind == 0) return write_double(&d, (const double *)values, file);
    if (kind == 1) return write_int(&d, (const int *)values, file);
    return write_float(&d, (const float *)values, file);
}

// one budgets CSV line (header first when iteration == start_iter)
int lemio_write_summary(const char *file, int iteration, int start_iter, double temp, double rain,
                        const double *tot8, int sinkcount)
{
    Data d;
    memset(&d, 0, sizeof d);
    snprintf(d.outfilename, sizeof d.outfilename, "%s", file);
    d.start_iter = start_iter;
    d.temp = temp;
    d.rain = rain;
    d.totE = tot8[0];
    d.totI = tot8[1];
    d.totG = tot8[2];
    d.totweatherC = tot8[3];
    d.totweatherP = tot8[4];
    d.totD = tot8[5];
    d.totbio = tot8[6];
    d.FA_max = tot8[7];
    d.sinkcount = sinkcount;
    return writeSummaryDataToFile(&d, iteration);
}

// climate series: n rows into year/rain/temp
int lemio_read_climate(const char *file, int n, int *year, double *rain, double *temp)
{
    Data d;
    memset(&d, 0, sizeof d);
    snprintf(d.clim_file, sizeof d.clim_file, "%s", file);
    d.max_iterations = n;
    int rc = ReadClimateDataFromFile(&d);
    if (rc == 0) {
        memcpy(year, d.year, n * sizeof(int));
        memcpy(rain, d.rainchange, n * sizeof(double));
        memcpy(temp, d.tempchange, n * sizeof(double));
    }
    free(d.year);
    free(d.rainchange);
    free(d.tempchange);
    return rc;
}

}  // extern "C"
