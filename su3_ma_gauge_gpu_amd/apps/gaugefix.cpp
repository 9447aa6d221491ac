// gaugefix.cpp -- the reference applications' control flow on the drop-in
// headers: LandauGaugeFixingSU3_4D.cu:62-286, U1xU1GaugeFixingSU3_4D.cu
// (its simulated-annealing schedule), MAGaugeFixingSU3_4D.cu:130-379 and
// CoulombGaugeFixingSU3_4D.cu:137-317.  Per gauge copy: optional random gauge
// transformation, simulated annealing (heatbath + microcanonical steps), the
// MAG stochastic-relaxation loop, the OR loop with reprojection every
// `reproject` sweeps and a quality check every `checkprecision` sweeps that
// stops below `precision`, a final projectSU3 and best-copy selection by gff.
//
// Lattice extents are compile-time constants, as in the reference
// (-D_X_=.. -D_T_=.. -DDOUBLEPRECISION).  Configurations are read from files
// (--fbasename given), made by setHot (--hotgaugefield 1), or else come from
// the library's synthetic generator (--sigma/--cfgseed/--haar).
//
// Options keep the reference's names and defaults
// (apps/program_options/ProgramOptions.hxx:229-266):
//   --gauge landau|u1xu1|mag|coulomb  --gaugecopies 1  --randomtrafo 1
//   --reproject 100  --seed 1  --sasteps 1000 --samin 0.01 --samax 0.4
//   --microupdates 3  --ormaxiter 1000 --orparameter 1.7  --srmaxiter 1000
//   --srparameter 1.7  --precision 1e-7  --checkprecision 100  --doSA 1
//   --hotgaugefield 0
//   files: --ftype PLAIN|HEADERONLY|VOGT|ILDG|QCDSTAG  --fbasename <path>
//   --fending .vogt  --fnumberformat 1  --fstartnumber 0  --fstepnumber 1
//   --nconf 1  --fappendix gaugefixed_  --reinterpret STANDARD|DOUBLE|FLOAT
//   --output_conf <path>  --output_ending ""  (MAG: --output_SA_functional
//   <path> --save_each 0).  File i is <fbasename><i><fending> with i padded
//   to fnumberformat digits, its output <output_conf><fappendix><i><output_ending>
//   (program_options/FileIterator.hxx:45-95).
//   Also --devicenumber/-D, -m for --nconf, --help, and a config file of
//   "key = value" lines (positional, or --config-file); the command line
//   overrides it (ProgramOptions.hxx:221-286).
// The last line printed per copy is "RESULT gff=<gff> theta=<theta> or_sweeps=<n> ..."
// (one per timeslice for Coulomb).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "culgt/CommonKernelsSU3.hxx"
#include "culgt/CoulombKernelsSU3.hxx"
#include "culgt/GaugeFixingStats.hxx"
#include "culgt/LandauKernelsSU3.hxx"
#include "culgt/MAGKernelsSU3.hxx"
#include "culgt/PhiloxWrapper.hxx"
#include "culgt/SiteIndex.hxx"
#include "culgt/U1xU1KernelsSU3.hxx"
#include "culgt/FileHeaderOnly.hxx"
#include "culgt/FilePlain.hxx"
#include "culgt/FileVogt.hxx"
#include "culgt/GpuPattern.hxx"
#include "culgt/LinkFile.hxx"
#include "culgt/SiteCoord.hxx"
#include "culgt/filetype_typedefs.h"
#include "culgt/ildg.hxx"
#include "culgt/qcdstag.hxx"

namespace {

struct Options {
    std::string gauge = "landau";
    int gaugecopies = 1, randomtrafo = 1, reproject = 100, seed = 1, sasteps = 1000, microupdates = 3;
    int ormaxiter = 1000, srmaxiter = 1000, checkprecision = 100, doSA = 1, hot = 0;
    float samin = 0.01f, samax = 0.4f, orparameter = 1.7f, srparameter = 1.7f, precision = 1e-7f;
    double sigma = 0.3;
    unsigned cfgseed = 1, haar = 2;
    // files (ProgramOptions.hxx:229-250)
    FileType ftype = PLAIN;
    bool ftype_set = false;
    ReinterpretReal reinterpret = STANDARD;
    std::string fbasename, fending = ".vogt", fappendix = "gaugefixed_", output_conf, output_ending;
    std::string output_SA_functional;
    int fnumberformat = 1, fstartnumber = 0, fstepnumber = 1, nconf = 1, save_each = 0;
    int devicenumber = -1;  // -1: the runtime's default device
};

std::string trim(const std::string &t)
{
    const size_t a = t.find_first_not_of(" \t\r"), b = t.find_last_not_of(" \t\r");
    return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
}

// "key = value" lines, '#' starting a comment (the config-file format of
// boost::program_options::parse_config_file used by ProgramOptions.hxx:258-262)
void read_config_file(const std::string &path, std::map<std::string, std::string> &kv)
{
    FILE *f = fopen(path.c_str(), "r");
    if (!f) {
        fprintf(stderr, "cannot open config file %s\n", path.c_str());
        exit(2);
    }
    char line[4096];
    while (fgets(line, sizeof line, f)) {
        std::string l(line);
        const size_t hash = l.find('#');
        if (hash != std::string::npos) l.resize(hash);
        const size_t eq = l.find('=');
        if (eq == std::string::npos) continue;
        const std::string k = trim(l.substr(0, eq));
        if (!k.empty() && !kv.count(k)) kv[k] = trim(l.substr(eq + 1));  // the command line wins
    }
    fclose(f);
}

const char *const kHelp =
    "Usage: gaugefix [options] [config-file]\n"
    "  --gauge landau|u1xu1|mag|coulomb, and the reference apps' options:\n"
    "  --config-file --devicenumber,-D --ftype --fbasename --fending --fnumberformat\n"
    "  --fstartnumber --fstepnumber --nconf,-m --fappendix --reinterpret --output_conf\n"
    "  --output_ending --output_SA_functional --save_each --doSA --hotgaugefield --seed\n"
    "  --gaugecopies --randomtrafo --reproject --sasteps --samin --samax --microupdates\n"
    "  --ormaxiter --orparameter --srmaxiter --srparameter --precision --checkprecision\n"
    "  (--sigma --cfgseed --haar: the generated configuration when no file is read)\n";

Options parse(int argc, char **argv)
{
    Options o;
    std::map<std::string, std::string> kv;
    std::string config;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--help") {
            printf("%s", kHelp);
            exit(1);
        }
        std::string key;
        if (a.rfind("--", 0) == 0) key = a.substr(2);
        else if (a == "-D") key = "devicenumber";
        else if (a == "-m") key = "nconf";
        if (key.empty()) {  // the positional argument is the config file
            config = a;
            continue;
        }
        if (i + 1 >= argc) {
            fprintf(stderr, "option %s needs a value\n", a.c_str());
            exit(2);
        }
        kv[key] = argv[++i];
    }
    if (kv.count("config-file")) config = kv["config-file"];
    if (!config.empty()) read_config_file(config, kv);
    o.devicenumber = kv.count("devicenumber") ? atoi(kv["devicenumber"].c_str()) : -1;
    auto I = [&](const char *k, int &v) { if (kv.count(k)) v = atoi(kv[k].c_str()); };
    auto F = [&](const char *k, float &v) { if (kv.count(k)) v = (float)atof(kv[k].c_str()); };
    if (kv.count("gauge")) o.gauge = kv["gauge"];
    I("gaugecopies", o.gaugecopies); I("randomtrafo", o.randomtrafo); I("reproject", o.reproject);
    I("seed", o.seed); I("sasteps", o.sasteps); I("microupdates", o.microupdates); I("ormaxiter", o.ormaxiter);
    I("srmaxiter", o.srmaxiter); I("checkprecision", o.checkprecision); I("doSA", o.doSA);
    I("hotgaugefield", o.hot);
    F("samin", o.samin); F("samax", o.samax); F("orparameter", o.orparameter); F("srparameter", o.srparameter);
    F("precision", o.precision);
    if (kv.count("sigma")) o.sigma = atof(kv["sigma"].c_str());
    if (kv.count("cfgseed")) o.cfgseed = (unsigned)atol(kv["cfgseed"].c_str());
    if (kv.count("haar")) o.haar = (unsigned)atol(kv["haar"].c_str());
    auto S = [&](const char *k, std::string &v) { if (kv.count(k)) v = kv[k]; };
    S("fbasename", o.fbasename); S("fending", o.fending); S("fappendix", o.fappendix);
    S("output_conf", o.output_conf); S("output_ending", o.output_ending);
    S("output_SA_functional", o.output_SA_functional);
    I("fnumberformat", o.fnumberformat); I("fstartnumber", o.fstartnumber); I("fstepnumber", o.fstepnumber);
    I("nconf", o.nconf); I("save_each", o.save_each);
    if (kv.count("ftype")) {
        std::istringstream in(kv["ftype"]);
        in >> o.ftype;
        o.ftype_set = true;
    }
    if (kv.count("reinterpret")) {
        std::istringstream in(kv["reinterpret"]);
        in >> o.reinterpret;
    }
    return o;
}

// program_options/FileIterator.hxx:45-95: file number `id` padded to
// fnumberformat digits
std::string padded(int id, int width)
{
    std::ostringstream s;
    s.width(width);
    s.fill('0');
    s << id;
    return s.str();
}

typedef StandardPattern<SiteCoord<4, NO_SPLIT>, 4, 3> Standard;

// The apps' file handling (LandauGaugeFixingSU3_4D.cu:118-121,147-183,296-323):
// one LinkFile per type, kept across load and save so that a HEADERONLY or
// VOGT header goes back out with the gauge-fixed links; ILDG and QCDSTAG
// through their functions.  Mem is GpuPattern (4D gauges) or
// GpuPatternTimeslice (Coulomb).
template <class Mem> struct ConfigFiles {
    LinkFile<FileHeaderOnly, Standard, Mem, SiteCoord<4, FULL_SPLIT> > headeronly;
    LinkFile<FileVogt, Standard, Mem, SiteCoord<4, FULL_SPLIT> > vogt;
    LinkFile<FilePlain, Standard, Mem, SiteCoord<4, FULL_SPLIT> > plain;
    explicit ConfigFiles(ReinterpretReal r) : headeronly(r), vogt(r), plain(r) {}
    static const bool slices = Mem::memory_pattern == GF_PATTERN_GPU_TIMESLICE;

    bool load(FileType t, const std::string &name, Real *U)
    {
        SiteCoord<4, FULL_SPLIT> s(HOST_CONSTANTS::SIZE);
        switch (t) {
        case VOGT: return vogt.load(s, name, U);
        case PLAIN: return plain.load(s, name, U);
        case HEADERONLY: return headeronly.load(s, name, U);
        case ILDG:
            if (slices) readILDG_timeslice(s, name.c_str(), HOST_CONSTANTS::SIZE, U);
            else readILDG(s, name.c_str(), HOST_CONSTANTS::SIZE, U);
            return true;
        case QCDSTAG:
            return slices ? readQCDSTAG_timeslice(s, name.c_str(), HOST_CONSTANTS::SIZE, U)
                          : readQCDSTAG(s, name.c_str(), HOST_CONSTANTS::SIZE, U);
        }
        return false;
    }

    bool save(FileType t, const std::string &in, const std::string &out, Real *U, int steps)
    {
        SiteCoord<4, FULL_SPLIT> s(HOST_CONSTANTS::SIZE);
        switch (t) {
        case VOGT: return vogt.save(s, out, U);
        case PLAIN: return plain.save(s, out, U);
        case HEADERONLY: return headeronly.save(s, out, U);
        case ILDG:
            if (slices) writeILDG_timeslice(s, in.c_str(), out.c_str(), HOST_CONSTANTS::SIZE, U, steps);
            else writeILDG(s, in.c_str(), out.c_str(), HOST_CONSTANTS::SIZE, U, steps);
            return true;
        case QCDSTAG:
            return slices ? writeQCDSTAG_timeslice(s, out.c_str(), HOST_CONSTANTS::SIZE, U)
                          : writeQCDSTAG(s, out.c_str(), HOST_CONSTANTS::SIZE, U);
        }
        return false;
    }
};

// what a run does besides fixing: MAG's per-copy functional file and the
// saving of each better copy (MAGaugeFixingSU3_4D.cu:151-154,336,347-374)
struct RunHooks {
    FILE *functional = nullptr;
    std::function<void(int copy, Real *U)> save_better;
};

#define HIPCHECK(x)                                                                  \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            exit(1);                                                                 \
        }                                                                            \
    } while (0)

const int64_t V = (int64_t)Nt * Nx * Ny * Nz;
const int64_t V3 = (int64_t)Nx * Ny * Nz;

// The host configuration in Real precision (the generator writes FP64).
std::vector<Real> make_config(const Options &o, const int dims[4], int64_t vol)
{
    std::vector<double> d((size_t)vol * 72);
    culgt_dropin::check(gf_make_config(dims, o.sigma, o.cfgseed, o.haar, d.data()));
    return std::vector<Real>(d.begin(), d.end());
}

// ---------------------------------------------------------------- 4D gauges
// LandauGaugeFixingSU3_4D.cu:184-286 (Landau), U1xU1GaugeFixingSU3_4D.cu (U1xU1 SA
// schedule), MAGaugeFixingSU3_4D.cu:199-345 (MAG).
// U: the host configuration (GpuPattern); on return it holds the best copy.
template <class GType> int run4d(const Options &o, std::vector<Real> &U, const RunHooks &hooks)
{
    const int threadsPerBlock = NSB * 8;
    const int numBlocks = (int)(V / 2 / NSB);
    const size_t bytes = (size_t)V * 72 * sizeof(Real);
    Real *dU;
    HIPCHECK(hipMalloc(&dU, bytes));
    // the reference's neighbour table (LandauGaugeFixingSU3_4D.cu:107-116);
    // the drop-in kernels accept it and compute neighbours on the fly
    SiteIndex<4, FULL_SPLIT> s(HOST_CONSTANTS::SIZE);
    std::vector<lat_index_t> nn((size_t)s.getLatticeSize() * 8);
    s.calculateNeighbourTable(nn.data());
    lat_index_t *dNn;
    HIPCHECK(hipMalloc(&dNn, nn.size() * sizeof(lat_index_t)));
    HIPCHECK(hipMemcpy(dNn, nn.data(), nn.size() * sizeof(lat_index_t), hipMemcpyHostToDevice));
    const bool mag = GType::gauge == GF_MAG, u1 = GType::gauge == GF_U1XU1;
    GaugeFixingStats<4, 3, GType, AVERAGE> gaugeStats(dU, HOST_CONSTANTS::SIZE);
    if (o.hot)
        CommonKernelsSU3::setHot((int)(V / 32), 32, dU, HOST_CONSTANTS::getPtrToDeviceSize(), o.seed,
                                 PhiloxWrapper::getNextCounter());
    double bestGff = 0.0;
    long orTotal = 0;
    for (int copy = 0; copy < o.gaugecopies; copy++) {
        if (!o.hot) HIPCHECK(hipMemcpy(dU, U.data(), bytes, hipMemcpyHostToDevice));
        if (mag) {
            gaugeStats.generateGaugeQuality();
            printf("initial functional %1.10f\n", gaugeStats.getCurrentGff());
        }
        if (o.randomtrafo) {
            GType::randomTrafo(numBlocks, threadsPerBlock, dU, dNn, 0, o.seed, PhiloxWrapper::getNextCounter());
            GType::randomTrafo(numBlocks, threadsPerBlock, dU, dNn, 1, o.seed, PhiloxWrapper::getNextCounter());
        }
        printf("i:\t\tgff:\t\tdA:\n");
        gaugeStats.generateGaugeQuality();
        printf("   \t\t%1.10f\t\t%e\n", gaugeStats.getCurrentGff(), gaugeStats.getCurrentA());

        // SIMULATED ANNEALING (Landau: sasteps steps; U1xU1: always, down to samin;
        // MAG: only with --doSA 1, five heatbath steps per temperature)
        if (!mag || o.doSA) {
            printf("SIMULATED ANNEALING\n");
            float temperature = o.samax;
            const float base = (o.samax - o.samin) / (float)o.sasteps;
            float tempStep = base;
            auto sa_sweep = [&]() {
                GType::saStep(numBlocks, threadsPerBlock, dU, dNn, 0, temperature, o.seed, PhiloxWrapper::getNextCounter());
                GType::saStep(numBlocks, threadsPerBlock, dU, dNn, 1, temperature, o.seed, PhiloxWrapper::getNextCounter());
                for (int mic = 0; mic < o.microupdates; mic++) {
                    GType::microStep(numBlocks, threadsPerBlock, dU, dNn, 0);
                    GType::microStep(numBlocks, threadsPerBlock, dU, dNn, 1);
                }
            };
            if (!mag && !u1) {
                for (int i = 0; i < o.sasteps; i++) {
                    sa_sweep();
                    if (i % o.reproject == 0)
                        CommonKernelsSU3::projectSU3((int)(V / 32), 32, dU, HOST_CONSTANTS::getPtrToDeviceSize());
                    if (i % o.checkprecision == 0) {
                        gaugeStats.generateGaugeQuality();
                        printf("%d\t\t%1.10f\t\t%e\n", i, gaugeStats.getCurrentGff(), gaugeStats.getCurrentA());
                    }
                    temperature -= tempStep;
                }
            } else {
                int i = 0;
                do {
                    // float temperature against double bounds, as the reference compares
                    if (u1) tempStep = (temperature <= 2.7 && temperature >= 2) ? base / 15 : base;
                    else tempStep = (temperature <= 0.88 && temperature >= 0.7) ? base / 15 : base;
                    if (mag) for (int j = 0; j < 5; j++) sa_sweep();
                    else sa_sweep();
                    if (i % o.reproject == 0)
                        CommonKernelsSU3::projectSU3((int)(V / 32), 32, dU, HOST_CONSTANTS::getPtrToDeviceSize());
                    if (u1 && i % o.checkprecision == 0) {
                        gaugeStats.generateGaugeQuality();
                        printf("%d\t\t%1.10f\t\t%e\n", i, gaugeStats.getCurrentGff(), gaugeStats.getCurrentA());
                    }
                    temperature -= tempStep;
                    i++;
                } while (temperature >= o.samin);
            }
        }

        // STOCHASTIC RELAXATION (MAG)
        int srSteps = 0;
        if (mag) {
            printf("STOCHASTIC RELAXATION\n");
            for (int i = 0; i < o.srmaxiter; i++) {
                MAGKernelsSU3::srStep(numBlocks, threadsPerBlock, dU, dNn, 0, o.srparameter, o.seed,
                                      PhiloxWrapper::getNextCounter());
                MAGKernelsSU3::srStep(numBlocks, threadsPerBlock, dU, dNn, 1, o.srparameter, o.seed,
                                      PhiloxWrapper::getNextCounter());
                srSteps = i + 1;
                if (i % o.reproject == 0)
                    CommonKernelsSU3::projectSU3((int)(V / 32), 32, dU, HOST_CONSTANTS::getPtrToDeviceSize());
                if (i % o.checkprecision == 0) {
                    gaugeStats.generateGaugeQuality();
                    printf("%d\t\t%1.10f\t\t%e\n", i, gaugeStats.getCurrentGff(), gaugeStats.getCurrentA());
                    if (gaugeStats.getCurrentA() < o.precision || gaugeStats.getCurrentA() < 1E-6) break;
                }
            }
        }

        // OVERRELAXATION
        printf("OVERRELAXATION\n");
        int orSteps = 0;
        for (int i = 0; i < o.ormaxiter; i++) {
            GType::orStep(numBlocks, threadsPerBlock, dU, dNn, 0, o.orparameter);
            GType::orStep(numBlocks, threadsPerBlock, dU, dNn, 1, o.orparameter);
            orSteps = i + 1;
            if (i % o.reproject == 0)
                CommonKernelsSU3::projectSU3((int)(V / 32), 32, dU, HOST_CONSTANTS::getPtrToDeviceSize());
            if (i % o.checkprecision == 0) {
                gaugeStats.generateGaugeQuality();
                printf("%d\t\t%1.10f\t\t%e\n", i, gaugeStats.getCurrentGff(), gaugeStats.getCurrentA());
                if (gaugeStats.getCurrentA() < o.precision) break;
            }
            orTotal++;
        }
        // reconstruct third line
        CommonKernelsSU3::projectSU3((int)(V / 32), 32, dU, HOST_CONSTANTS::getPtrToDeviceSize());
        if (mag) {
            gaugeStats.generateGaugeQuality();
            printf("%d,%.17g\n", copy, gaugeStats.getCurrentGff());
            if (hooks.functional) fprintf(hooks.functional, "%d,%.17g\n", copy, gaugeStats.getCurrentGff());
        }
        const bool better = gaugeStats.getCurrentGff() > bestGff;
        if (better) {
            printf("FOUND BETTER COPY\n");
            bestGff = gaugeStats.getCurrentGff();
            HIPCHECK(hipMemcpy(U.data(), dU, bytes, hipMemcpyDeviceToHost));
            if (hooks.save_better && copy < o.gaugecopies - 1) hooks.save_better(copy, U.data());
        } else {
            printf("NO BETTER COPY\n");
        }
        printf("RESULT copy=%d gff=%.17g theta=%.17g or_sweeps=%d sr_sweeps=%d counter=%u\n", copy,
               gaugeStats.getCurrentGff(), gaugeStats.getCurrentA(), orSteps, srSteps,
               PhiloxWrapper::getCurrentCounter());
    }
    printf("BEST gff=%.17g\n", bestGff);
    HIPCHECK(hipFree(dU));
    HIPCHECK(hipFree(dNn));
    return 0;
}

// ------------------------------------------------------------------ Coulomb
// CoulombGaugeFixingSU3_4D.cu:137-317: one timeslice after the other; slice t
// and slice t-1 are copied to the device, fixed, and written back if better.
// The host array is in the timeslice layout (GpuPatternTimeslice.hxx:57-61).
// 4D GpuPattern array -> the timeslice layout (GpuPatternTimeslice.hxx:57-61)
std::vector<Real> to_timeslices(const std::vector<Real> &U4)
{
    const int64_t tsSize = V3 * 72;
    std::vector<Real> U((size_t)V * 72);
    for (int t = 0; t < Nt; t++)
        for (int x = 0; x < Nx; x++)
            for (int y = 0; y < Ny; y++)
                for (int z = 0; z < Nz; z++) {
                    const int64_t n4 = (((int64_t)t * Nx + x) * Ny + y) * Nz + z;
                    const int64_t i4 = ((t + x + y + z) % 2 == 0) ? n4 / 2 : n4 / 2 + V / 2;
                    const int64_t n3 = ((int64_t)x * Ny + y) * Nz + z;
                    const int64_t i3 = ((x + y + z) % 2 == 0) ? n3 / 2 : n3 / 2 + V3 / 2;
                    for (int k = 0; k < 72; k++) U[(size_t)(t * tsSize + i3 + V3 * k)] = U4[(size_t)(i4 + V * k)];
                }
    return U;
}

// U: the host configuration in the timeslice layout; on return each slice
// holds its best copy.
int run_coulomb(const Options &o, std::vector<Real> &U)
{
    const int threadsPerBlock = NSB * 8;
    const int numBlocks = (int)(V3 / 2 / NSB);
    const int64_t tsSize = V3 * 72;  // reals per timeslice
    const size_t bytes = (size_t)tsSize * sizeof(Real);
    Real *dUtUp, *dUtDw;
    HIPCHECK(hipMalloc(&dUtUp, bytes));
    HIPCHECK(hipMalloc(&dUtDw, bytes));
    lat_index_t *dNnt = nullptr;
    GaugeFixingStats<4, 3, CoulombKernelsSU3, AVERAGE> gaugeStats(dUtUp, HOST_CONSTANTS::SIZE_TIMESLICE);
    for (int t = 0; t < Nt; t++) {
        const int tDw = (t > 0) ? (t - 1) : (Nt - 1);
        double bestGff = 0.0;
        for (int copy = 0; copy < o.gaugecopies; copy++) {
            if (!o.hot) {
                HIPCHECK(hipMemcpy(dUtUp, &U[(size_t)t * tsSize], bytes, hipMemcpyHostToDevice));
                HIPCHECK(hipMemcpy(dUtDw, &U[(size_t)tDw * tsSize], bytes, hipMemcpyHostToDevice));
            } else {
                CommonKernelsSU3::setHot((int)(V3 / 32), 32, dUtUp, HOST_CONSTANTS::getPtrToDeviceSizeTimeslice(), o.seed,
                                         PhiloxWrapper::getNextCounter());
                CommonKernelsSU3::setHot((int)(V3 / 32), 32, dUtDw, HOST_CONSTANTS::getPtrToDeviceSizeTimeslice(), o.seed,
                                         PhiloxWrapper::getNextCounter());
            }
            if (o.randomtrafo) {
                CoulombKernelsSU3::randomTrafo(numBlocks, threadsPerBlock, dUtUp, dUtDw, dNnt, 0, o.seed,
                                               PhiloxWrapper::getNextCounter());
                CoulombKernelsSU3::randomTrafo(numBlocks, threadsPerBlock, dUtUp, dUtDw, dNnt, 1, o.seed,
                                               PhiloxWrapper::getNextCounter());
            }
            gaugeStats.generateGaugeQuality();
            float temperature = o.samax;
            const float tempStep = (o.samax - o.samin) / (float)o.sasteps;
            for (int i = 0; i < o.sasteps; i++) {
                CoulombKernelsSU3::saStep(numBlocks, threadsPerBlock, dUtUp, dUtDw, dNnt, 0, temperature, o.seed,
                                          PhiloxWrapper::getNextCounter());
                CoulombKernelsSU3::saStep(numBlocks, threadsPerBlock, dUtUp, dUtDw, dNnt, 1, temperature, o.seed,
                                          PhiloxWrapper::getNextCounter());
                for (int mic = 0; mic < o.microupdates; mic++) {
                    CoulombKernelsSU3::microStep(numBlocks, threadsPerBlock, dUtUp, dUtDw, dNnt, 0);
                    CoulombKernelsSU3::microStep(numBlocks, threadsPerBlock, dUtUp, dUtDw, dNnt, 1);
                }
                if (i % o.reproject == 0) {
                    CommonKernelsSU3::projectSU3((int)(V3 / 32), 32, dUtUp, HOST_CONSTANTS::getPtrToDeviceSizeTimeslice());
                    CommonKernelsSU3::projectSU3((int)(V3 / 32), 32, dUtDw, HOST_CONSTANTS::getPtrToDeviceSizeTimeslice());
                }
                if (i % o.checkprecision == 0) gaugeStats.generateGaugeQuality();
                temperature -= tempStep;
            }
            int orSteps = 0;
            for (int i = 0; i < o.ormaxiter; i++) {
                CoulombKernelsSU3::orStep(numBlocks, threadsPerBlock, dUtUp, dUtDw, dNnt, 0, o.orparameter);
                CoulombKernelsSU3::orStep(numBlocks, threadsPerBlock, dUtUp, dUtDw, dNnt, 1, o.orparameter);
                orSteps = i + 1;
                if (i % o.reproject == 0) {
                    CommonKernelsSU3::projectSU3((int)(V3 / 32), 32, dUtUp, HOST_CONSTANTS::getPtrToDeviceSizeTimeslice());
                    CommonKernelsSU3::projectSU3((int)(V3 / 32), 32, dUtDw, HOST_CONSTANTS::getPtrToDeviceSizeTimeslice());
                }
                if (i % o.checkprecision == 0) {
                    gaugeStats.generateGaugeQuality();
                    if (gaugeStats.getCurrentA() < o.precision) break;
                }
            }
            CommonKernelsSU3::projectSU3((int)(V3 / 32), 32, dUtUp, HOST_CONSTANTS::getPtrToDeviceSizeTimeslice());
            CommonKernelsSU3::projectSU3((int)(V3 / 32), 32, dUtDw, HOST_CONSTANTS::getPtrToDeviceSizeTimeslice());
            printf("RESULT slice=%d copy=%d gff=%.17g theta=%.17g or_sweeps=%d counter=%u\n", t, copy,
                   gaugeStats.getCurrentGff(), gaugeStats.getCurrentA(), orSteps, PhiloxWrapper::getCurrentCounter());
            if (gaugeStats.getCurrentGff() > bestGff) {
                bestGff = gaugeStats.getCurrentGff();
                HIPCHECK(hipMemcpy(&U[(size_t)t * tsSize], dUtUp, bytes, hipMemcpyDeviceToHost));
                HIPCHECK(hipMemcpy(&U[(size_t)tDw * tsSize], dUtDw, bytes, hipMemcpyDeviceToHost));
            }
        }
    }
    HIPCHECK(hipFree(dUtUp));
    HIPCHECK(hipFree(dUtDw));
    return 0;
}

int run_one(const Options &o, std::vector<Real> &U, const RunHooks &hooks)
{
    if (o.gauge == "landau") return run4d<LandauKernelsSU3>(o, U, hooks);
    if (o.gauge == "u1xu1") return run4d<U1xU1KernelsSU3>(o, U, hooks);
    if (o.gauge == "mag") return run4d<MAGKernelsSU3>(o, U, hooks);
    return run_coulomb(o, U);
}

// every file of the FileIterator: load, fix, save (the apps' main loops)
template <class Mem> int run_files(const Options &o)
{
    ConfigFiles<Mem> files(o.reinterpret);
    const bool mag = o.gauge == "mag";
    for (int id = o.fstartnumber; id < o.fstartnumber + o.nconf * o.fstepnumber; id += o.fstepnumber) {
        const std::string in = o.fbasename + padded(id, o.fnumberformat) + o.fending;
        const std::string out = o.output_conf + o.fappendix + padded(id, o.fnumberformat) + o.output_ending;
        std::vector<Real> U((size_t)V * 72);
        printf("loading %s\n", in.c_str());
        if (!files.load(o.ftype, in, U.data())) {
            printf("Error while loading. Trying next file.\n");
            break;  // the reference leaves the loop here
        }
        printf("File loaded.\n");
        RunHooks hooks;
        FILE *functional = nullptr;
        if (mag && !o.output_SA_functional.empty()) {
            const std::string f = o.output_SA_functional + "_" + padded(id, o.fnumberformat);
            functional = fopen(f.c_str(), "w");
            if (functional) fprintf(functional, "copy,functional\n");
            hooks.functional = functional;
        }
        if (mag && o.save_each)
            hooks.save_better = [&](int copy, Real *Ub) {
                files.save(o.ftype, in, out + "_" + std::to_string(copy + 1), Ub, o.sasteps);
            };
        const int rc = run_one(o, U, hooks);
        if (functional) fclose(functional);
        if (rc != 0) return rc;
        if (!(mag && o.save_each)) {
            printf("saving %s\n", out.c_str());
            if (!files.save(o.ftype, in, out, U.data(), o.sasteps)) return 1;
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    const Options o = parse(argc, argv);
    if (o.devicenumber >= 0) HIPCHECK(hipSetDevice(o.devicenumber));
    if (o.gauge != "landau" && o.gauge != "u1xu1" && o.gauge != "mag" && o.gauge != "coulomb") {
        fprintf(stderr, "unknown gauge %s\n", o.gauge.c_str());
        return 2;
    }
    try {
        if (!o.hot && !o.fbasename.empty()) {
            if (!o.ftype_set) {
                fprintf(stderr, "Filetype not set to a known value. Exiting...\n");
                return 1;
            }
            if (o.gauge == "coulomb") return run_files<GpuPatternTimeslice<SiteCoord<4, FULL_SPLIT>, 4, 3> >(o);
            return run_files<GpuPattern<SiteCoord<4, FULL_SPLIT>, 4, 3> >(o);
        }
        const int dims[4] = {Nt, Nx, Ny, Nz};
        std::vector<Real> U = o.hot ? std::vector<Real>((size_t)V * 72) : make_config(o, dims, V);
        if (o.gauge == "coulomb" && !o.hot) U = to_timeslices(U);
        return run_one(o, U, RunHooks());
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
