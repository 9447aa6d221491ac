// ref_io_harness.cpp -- exposes the reference's configuration-file code
// through a C ABI so that the library's file formats can be checked against it.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target "ref") from the
// reference sources where they lie; the output goes to oracle/_ref/ and is
// loaded only by the test-suite.
//
// Compiled from the reference (paths relative to its src/):
//   gaugefixing/apps/ildg.cpp      readILDG, writeILDG, readILDG_timeslice, writeILDG_timeslice
//   gaugefixing/apps/qcdstag.cpp   readQCDSTAG, writeQCDSTAG and the _timeslice forms
//   gaugefixing/apps/src/c-lime/*.c  the LIME reader and writer they call
//   lattice/filetypes/FileVogt.hxx   the VOGT header (included here)
//   lattice/SiteCoord.hxx, lattice/access_pattern/{GpuPattern,GpuPatternTimeslice,StandardPattern}.hxx
// These are compiled as their own translation units by the Makefile, as C++
// (the way gaugefixing/apps/Makefile:88-107 builds them).  lattice/LinkFile.hxx
// needs boost (filetype_typedefs.h) and FileHeaderOnly.hxx includes <cuda.h>
// through GlobalConstants.h: neither is buildable here.  The standard-order
// loop of LinkFile::load / save (LinkFile.hxx:119-150, 219-252) is therefore
// composed below from the reference's own pattern classes.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

#include "lattice/datatype/datatypes.h"
#include "lattice/SiteCoord.hxx"
#include "lattice/access_pattern/GpuPattern.hxx"
#include "lattice/access_pattern/GpuPatternTimeslice.hxx"
#include "lattice/access_pattern/StandardPattern.hxx"
#include "lattice/filetypes/FileVogt.hxx"
#include "gaugefixing/apps/include/c-lime/lime.h"

// compiled from gaugefixing/apps/ildg.cpp and qcdstag.cpp
void readILDG(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const short SIZE[4], Real *U);
void writeILDG(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const char *output_name, const short SIZE[4], Real *U,
               int steps);
void readILDG_timeslice(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const short SIZE[4], Real *U);
void writeILDG_timeslice(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const char *output_name,
                         const short SIZE[4], Real *U, int steps);
bool readQCDSTAG(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const short SIZE[4], Real *U);
bool writeQCDSTAG(SiteCoord<4, FULL_SPLIT> s, const char *output_name, const short SIZE[4], Real *U);
bool readQCDSTAG_timeslice(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const short SIZE[4], Real *U);
bool writeQCDSTAG_timeslice(SiteCoord<4, FULL_SPLIT> s, const char *output_name, const short SIZE[4], Real *U);

#define REF_EXPORT extern "C" __attribute__((visibility("default")))

typedef SiteCoord<4, FULL_SPLIT> RCoord;
typedef SiteCoord<4, NO_SPLIT> RCoordLex;
typedef StandardPattern<RCoordLex, 4, 3> RStandard;

static void sizes(const int dims[4], short S[4], lat_coord_t L[4])
{
    for (int i = 0; i < 4; i++) {
        S[i] = (short)dims[i];
        L[i] = (lat_coord_t)dims[i];
    }
}

REF_EXPORT int ref_io_real_size() { return (int)sizeof(Real); }

REF_EXPORT void ref_read_ildg(const int dims[4], const char *path, double *U, int timeslice)
{
    short S[4];
    lat_coord_t L[4];
    sizes(dims, S, L);
    RCoord s(L);
    if (timeslice) readILDG_timeslice(s, path, S, U);
    else readILDG(s, path, S, U);
}

REF_EXPORT void ref_write_ildg(const int dims[4], const char *tmpl, const char *out, double *U, int steps,
                               int timeslice)
{
    short S[4];
    lat_coord_t L[4];
    sizes(dims, S, L);
    RCoord s(L);
    if (timeslice) writeILDG_timeslice(s, tmpl, out, S, U, steps);
    else writeILDG(s, tmpl, out, S, U, steps);
}

REF_EXPORT int ref_read_qcdstag(const int dims[4], const char *path, double *U, int timeslice)
{
    short S[4];
    lat_coord_t L[4];
    sizes(dims, S, L);
    RCoord s(L);
    return timeslice ? readQCDSTAG_timeslice(s, path, S, U) : readQCDSTAG(s, path, S, U);
}

REF_EXPORT int ref_write_qcdstag(const int dims[4], const char *path, double *U, int timeslice)
{
    short S[4];
    lat_coord_t L[4];
    sizes(dims, S, L);
    RCoord s(L);
    return timeslice ? writeQCDSTAG_timeslice(s, path, S, U) : writeQCDSTAG(s, path, S, U);
}

// host-array index of real u of the standard order: MemoryPattern::getIndexByUnique(
// FilePattern::getUniqueIndex(u, size), size) as in LinkFile.hxx:125-126
REF_EXPORT long long ref_index_by_unique(const int dims[4], long long u, int timeslice)
{
    short S[4];
    lat_coord_t L[4];
    sizes(dims, S, L);
    const lat_array_index_t uq = RStandard::getUniqueIndex((lat_array_index_t)u, L);
    if (timeslice) return (long long)GpuPatternTimeslice<RCoord, 4, 3>::getIndexByUnique(uq, L);
    return (long long)GpuPattern<RCoord, 4, 3>::getIndexByUnique(uq, L);
}

// a VOGT file written by the reference's header code plus the standard-order
// array of LinkFile::save (LinkFile.hxx:183-252); header fields as given
REF_EXPORT int ref_write_vogt(const int dims[4], const short *header, int nheader_dims, const char *path,
                              const double *U)
{
    FileVogt fv(8);
    fv.ndim = (short)nheader_dims;
    fv.nc = header[0];
    for (int i = 0; i < nheader_dims; i++) fv.latsize[i] = header[1 + i];
    fv.lengthOfReal = header[1 + nheader_dims];
    std::fstream file;
    file.open(path, std::ios::out | std::ios::binary);
    if (!file) return 0;
    if (!fv.saveHeader(&file)) return 0;
    const long long n = 72LL * dims[0] * dims[1] * dims[2] * dims[3];
    for (long long u = 0; u < n; u++) {
        const double v = U[ref_index_by_unique(dims, u, 0)];
        file.write((const char *)&v, sizeof v);
    }
    file.close();
    return 1;
}

// the VOGT header fields the reference's loadHeader reads (FileVogt.hxx:52-79):
// out = {ndim, nc, latsize[0..ndim-1], lengthOfReal}
REF_EXPORT int ref_read_vogt_header(const char *path, short *out)
{
    FileVogt fv(8);
    std::fstream file;
    file.open(path, std::ios::in | std::ios::binary);
    if (!file) return 0;
    if (!fv.loadHeader(&file)) return 0;
    out[0] = fv.ndim;
    out[1] = fv.nc;
    for (int i = 0; i < fv.ndim && i < 16; i++) out[2 + i] = fv.latsize[i];
    out[2 + fv.ndim] = fv.lengthOfReal;
    return 1;
}

// write a LIME file with c-lime: n records of the given types, MB/ME flags and data
REF_EXPORT int ref_lime_write(const char *path, int n, const char **types, const int *mb, const int *me,
                              const char **data, const unsigned long long *nbytes)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return 0;
    LimeWriter *w = limeCreateWriter(fp);
    int ok = 1;
    for (int k = 0; k < n && ok; k++) {
        n_uint64_t len = nbytes[k];
        LimeRecordHeader *h = limeCreateHeader(mb[k], me[k], (char *)types[k], len);
        if (limeWriteRecordHeader(h, w) < 0) ok = 0;
        limeDestroyHeader(h);
        if (ok && limeWriteRecordData((void *)data[k], &len, w) != LIME_SUCCESS) ok = 0;
    }
    limeDestroyWriter(w);
    fclose(fp);
    return ok;
}

// list the records of a LIME file with c-lime: for record k, types[k] (up to
// 127 chars), flags[k] = 2 MB + ME, nbytes[k]; returns the number of records
// (at most max), or -1 on a reader error
REF_EXPORT int ref_lime_list(const char *path, int max, char *types, int *flags, unsigned long long *nbytes)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    LimeReader *r = limeCreateReader(fp);
    int k = 0, status;
    while ((status = limeReaderNextRecord(r)) != LIME_EOF) {
        if (status != LIME_SUCCESS) {
            k = -1;
            break;
        }
        if (k < max) {
            strncpy(types + 128 * k, limeReaderType(r), 127);
            types[128 * k + 127] = 0;
            flags[k] = 2 * limeReaderMBFlag(r) + limeReaderMEFlag(r);
            nbytes[k] = limeReaderBytes(r);
        }
        k++;
    }
    limeDestroyReader(r);
    fclose(fp);
    return k;
}

// the data of record k of a LIME file, read with c-lime (returns its size)
REF_EXPORT long long ref_lime_read(const char *path, int index, char *out, unsigned long long cap)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    LimeReader *r = limeCreateReader(fp);
    long long got = -1;
    int k = 0, status;
    while ((status = limeReaderNextRecord(r)) != LIME_EOF) {
        if (status != LIME_SUCCESS) break;
        if (k == index) {
            n_uint64_t n = limeReaderBytes(r);
            if (n > cap) n = cap;
            if (limeReaderReadData(out, &n, r) == LIME_SUCCESS) got = (long long)n;
            break;
        }
        k++;
    }
    limeDestroyReader(r);
    fclose(fp);
    return got;
}
