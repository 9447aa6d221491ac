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
// (-D_X_=.. -D_T_=.. -DDOUBLEPRECISION).  Configurations come from the
// library's synthetic generator (--sigma/--cfgseed/--haar) or setHot
// (--hotgaugefield 1); file input/output is not part of this build.
//
// Options keep the reference's names and defaults
// (apps/program_options/ProgramOptions.hxx:229-266):
//   --gauge landau|u1xu1|mag|coulomb  --gaugecopies 1  --randomtrafo 1
//   --reproject 100  --seed 1  --sasteps 1000 --samin 0.01 --samax 0.4
//   --microupdates 3  --ormaxiter 1000 --orparameter 1.7  --srmaxiter 1000
//   --srparameter 1.7  --precision 1e-7  --checkprecision 100  --doSA 1
//   --hotgaugefield 0   (plus --sigma 0.3 --cfgseed 1 --haar 2 for the input)
// The last line printed is "RESULT gff=<gff> theta=<theta> or_sweeps=<n> ..."
// (one per timeslice for Coulomb).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
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

namespace {

struct Options {
    std::string gauge = "landau";
    int gaugecopies = 1, randomtrafo = 1, reproject = 100, seed = 1, sasteps = 1000, microupdates = 3;
    int ormaxiter = 1000, srmaxiter = 1000, checkprecision = 100, doSA = 1, hot = 0;
    float samin = 0.01f, samax = 0.4f, orparameter = 1.7f, srparameter = 1.7f, precision = 1e-7f;
    double sigma = 0.3;
    unsigned cfgseed = 1, haar = 2;
};

Options parse(int argc, char **argv)
{
    Options o;
    std::map<std::string, std::string> kv;
    for (int i = 1; i + 1 < argc; i += 2) {
        if (strncmp(argv[i], "--", 2) != 0) {
            fprintf(stderr, "unexpected argument %s\n", argv[i]);
            exit(2);
        }
        kv[argv[i] + 2] = argv[i + 1];
    }
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
    return o;
}

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
template <class GType> int run4d(const Options &o)
{
    const int threadsPerBlock = NSB * 8;
    const int numBlocks = (int)(V / 2 / NSB);
    const size_t bytes = (size_t)V * 72 * sizeof(Real);
    const int dims[4] = {Nt, Nx, Ny, Nz};
    std::vector<Real> U = o.hot ? std::vector<Real>((size_t)V * 72) : make_config(o, dims, V);
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
        }
        const bool better = gaugeStats.getCurrentGff() > bestGff;
        if (better) {
            printf("FOUND BETTER COPY\n");
            bestGff = gaugeStats.getCurrentGff();
            HIPCHECK(hipMemcpy(U.data(), dU, bytes, hipMemcpyDeviceToHost));
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
int run_coulomb(const Options &o)
{
    const int threadsPerBlock = NSB * 8;
    const int numBlocks = (int)(V3 / 2 / NSB);
    const int64_t tsSize = V3 * 72;  // reals per timeslice
    const size_t bytes = (size_t)tsSize * sizeof(Real);
    const int dims[4] = {Nt, Nx, Ny, Nz};
    // 4D generator output -> timeslice layout
    std::vector<Real> U4 = o.hot ? std::vector<Real>((size_t)V * 72) : make_config(o, dims, V);
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

}  // namespace

int main(int argc, char **argv)
{
    const Options o = parse(argc, argv);
    try {
        if (o.gauge == "landau") return run4d<LandauKernelsSU3>(o);
        if (o.gauge == "u1xu1") return run4d<U1xU1KernelsSU3>(o);
        if (o.gauge == "mag") return run4d<MAGKernelsSU3>(o);
        if (o.gauge == "coulomb") return run_coulomb(o);
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    fprintf(stderr, "unknown gauge %s\n", o.gauge.c_str());
    return 2;
}
