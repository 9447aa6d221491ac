// culgt/qcdstag.hxx -- the QCDSTAG functions of the reference's apps
// (gaugefixing/apps/qcdstag.cpp:17-227), on the library's file code
// (include/su3gf_io.h).  They return false on failure, as the reference's do.
#ifndef CULGT_QCDSTAG_HXX_
#define CULGT_QCDSTAG_HXX_

#include "ildg.hxx"

namespace culgt_dropin {
inline bool qcdstag(bool save, const char *name, const short SIZE[4], Real *U, int pattern)
{
    FileHandle f(GF_FILE_QCDSTAG, pattern);
    const gf_lattice l = size_lattice(SIZE);
    if (save) return file_ok(gf_linkfile_save(f.h, name, &l, U, 0), "writeQCDSTAG error:", name);
    return file_ok(gf_linkfile_load(f.h, name, &l, U), "readQCDSTAG error:", name);
}
}  // namespace culgt_dropin

inline bool readQCDSTAG(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const short SIZE[4], Real *U)
{
    (void)s;
    return culgt_dropin::qcdstag(false, file_name, SIZE, U, GF_PATTERN_GPU);
}

inline bool writeQCDSTAG(SiteCoord<4, FULL_SPLIT> s, const char *output_name, const short SIZE[4], Real *U)
{
    (void)s;
    return culgt_dropin::qcdstag(true, output_name, SIZE, U, GF_PATTERN_GPU);
}

inline bool readQCDSTAG_timeslice(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const short SIZE[4], Real *U)
{
    (void)s;
    return culgt_dropin::qcdstag(false, file_name, SIZE, U, GF_PATTERN_GPU_TIMESLICE);
}

inline bool writeQCDSTAG_timeslice(SiteCoord<4, FULL_SPLIT> s, const char *output_name, const short SIZE[4],
                                   Real *U)
{
    (void)s;
    return culgt_dropin::qcdstag(true, output_name, SIZE, U, GF_PATTERN_GPU_TIMESLICE);
}

#endif
