// Exercises the drop-in file headers (culgt/LinkFile.hxx, ildg.hxx,
// qcdstag.hxx, GpuPattern.hxx, SiteCoord.hxx) the way the reference's apps
// call them.  Built by tests/test_culgt_files.py with -D_T_=2 -D_X_=4 -D_Y_=6
// -D_Z_=8 -DDOUBLEPRECISION.
//   linkfile_check index               -> the GpuPattern / GpuPatternTimeslice index of every unique index
//   linkfile_check write <in> <dir>    -> <in>: raw GpuPattern doubles; writes dir/{plain,vogt,qcdstag,qcdstag_ts}
//                                         and dir/ildg (from the template dir/template.lime, 5 SA steps)
//   linkfile_check read <dir>          -> reads them back, prints "ok <format>" for each exact match
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "culgt/FileHeaderOnly.hxx"
#include "culgt/FilePlain.hxx"
#include "culgt/FileVogt.hxx"
#include "culgt/GpuPattern.hxx"
#include "culgt/LinkFile.hxx"
#include "culgt/SiteCoord.hxx"
#include "culgt/filetype_typedefs.h"
#include "culgt/ildg.hxx"
#include "culgt/qcdstag.hxx"

typedef StandardPattern<SiteCoord<4, NO_SPLIT>, 4, 3> Standard;
typedef GpuPattern<SiteCoord<4, FULL_SPLIT>, 4, 3> Gpu;
typedef GpuPatternTimeslice<SiteCoord<4, FULL_SPLIT>, 4, 3> GpuTimeslice;

static const int arraySize = Nt * Nx * Ny * Nz * 4 * 3 * 3 * 2;

static std::vector<Real> raw(const std::string &path)
{
    std::vector<Real> U(arraySize);
    FILE *f = fopen(path.c_str(), "rb");
    if (!f || fread(U.data(), sizeof(Real), U.size(), f) != U.size()) {
        fprintf(stderr, "cannot read %s\n", path.c_str());
        exit(1);
    }
    fclose(f);
    return U;
}

int main(int argc, char **argv)
{
    SiteCoord<4, FULL_SPLIT> s(HOST_CONSTANTS::SIZE);
    lat_coord_t size[4] = {Nt, Nx, Ny, Nz};
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "index") {
        for (int u = 0; u < arraySize; u++)
            printf("%d %d\n", (int)Gpu::getIndexByUnique(u, size), (int)GpuTimeslice::getIndexByUnique(u, size));
        return 0;
    }
    if (mode == "write" && argc == 4) {
        std::vector<Real> U = raw(argv[2]);
        const std::string d = argv[3];
        LinkFile<FilePlain, Standard, Gpu, SiteCoord<4, FULL_SPLIT> > lfPlain(STANDARD);
        LinkFile<FileVogt, Standard, Gpu, SiteCoord<4, FULL_SPLIT> > lfVogt(STANDARD);
        if (!lfPlain.save(s, d + "/plain", U.data())) return 1;
        if (!lfVogt.save(s, d + "/vogt", U.data())) return 1;
        if (!writeQCDSTAG(s, (d + "/qcdstag").c_str(), HOST_CONSTANTS::SIZE, U.data())) return 1;
        if (!writeQCDSTAG_timeslice(s, (d + "/qcdstag_ts").c_str(), HOST_CONSTANTS::SIZE, U.data())) return 1;
        writeILDG(s, (d + "/template.lime").c_str(), (d + "/ildg").c_str(), HOST_CONSTANTS::SIZE, U.data(), 5);
        return 0;
    }
    if (mode == "read" && argc == 4) {
        std::vector<Real> U = raw(argv[2]);
        const std::string d = argv[3];
        std::vector<Real> W(arraySize);
        LinkFile<FilePlain, Standard, Gpu, SiteCoord<4, FULL_SPLIT> > lfPlain(STANDARD);
        LinkFile<FileVogt, Standard, Gpu, SiteCoord<4, FULL_SPLIT> > lfVogt(STANDARD);
        LinkFile<FileHeaderOnly, Standard, Gpu, SiteCoord<4, FULL_SPLIT> > lfHeaderOnly(STANDARD);
        if (lfPlain.load(s, d + "/plain", W.data()) && W == U) printf("ok plain\n");
        W.assign(arraySize, 0);
        if (lfVogt.load(s, d + "/vogt", W.data()) && W == U) printf("ok vogt\n");
        W.assign(arraySize, 0);
        if (lfHeaderOnly.load(s, d + "/headeronly", W.data()) && W == U) printf("ok headeronly\n");
        if (lfHeaderOnly.save(s, d + "/headeronly_out", W.data())) printf("saved headeronly\n");
        W.assign(arraySize, 0);
        if (readQCDSTAG(s, (d + "/qcdstag").c_str(), HOST_CONSTANTS::SIZE, W.data()) && W == U) printf("ok qcdstag\n");
        W.assign(arraySize, 0);
        readILDG(s, (d + "/ildg").c_str(), HOST_CONSTANTS::SIZE, W.data());
        if (W == U) printf("ok ildg\n");
        // a missing file: false and a message, as the reference's LinkFile
        if (!lfPlain.load(s, d + "/missing", W.data())) printf("ok missing\n");
        // the option parsers
        FileType ft = PLAIN;
        ReinterpretReal rr = STANDARD;
        std::string a = "ildg", b = "Float";
        std::istringstream ia(a), ib(b);
        ia >> ft;
        ib >> rr;
        if (ft == ILDG && rr == FLOAT) printf("ok options\n");
        return 0;
    }
    fprintf(stderr, "usage: linkfile_check index | write <raw> <dir> | read <raw> <dir>\n");
    return 2;
}
