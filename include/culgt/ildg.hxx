// culgt/ildg.hxx -- the ILDG functions of the reference's apps
// (gaugefixing/apps/ildg.cpp:26-451), on the library's LIME reader and writer
// (include/su3gf_io.h).  writeILDG copies the records of file_name (the file
// the configuration was read from) and replaces its binary data, appending
// " SA steps <steps>" to the xlf-info record, as the reference does.
#ifndef CULGT_ILDG_HXX_
#define CULGT_ILDG_HXX_

#include <string>

#include "LinkFile.hxx"

namespace culgt_dropin {
inline gf_lattice size_lattice(const short SIZE[4])
{
    gf_lattice l;
    for (int i = 0; i < 4; i++) l.dims[i] = SIZE[i];
    l.dtype = sizeof(Real) == sizeof(double) ? GF_F64 : GF_F32;
    l.layout = GF_LAYOUT_PATTERN18;
    return l;
}

inline void read_ildg(const char *file_name, const short SIZE[4], Real *U, int pattern)
{
    FileHandle f(GF_FILE_ILDG, pattern);
    const gf_lattice l = size_lattice(SIZE);
    file_ok(gf_linkfile_load(f.h, file_name, &l, U), "cannot read ILDG file", file_name);
}

inline void write_ildg(const char *file_name, const char *output_name, const short SIZE[4], Real *U, int steps,
                       int pattern)
{
    FileHandle f(GF_FILE_ILDG, pattern);
    const gf_lattice l = size_lattice(SIZE);
    if (!file_ok(gf_linkfile_read_header(f.h, file_name, &l), "cannot read ILDG records of", file_name)) return;
    file_ok(gf_linkfile_save(f.h, output_name, &l, U, steps), "cannot write ILDG file", output_name);
}
}  // namespace culgt_dropin

inline void readILDG(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const short SIZE[4], Real *U)
{
    (void)s;
    culgt_dropin::read_ildg(file_name, SIZE, U, GF_PATTERN_GPU);
}

inline void writeILDG(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const char *output_name, const short SIZE[4],
                      Real *U, int steps)
{
    (void)s;
    culgt_dropin::write_ildg(file_name, output_name, SIZE, U, steps, GF_PATTERN_GPU);
}

inline void readILDG_timeslice(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const short SIZE[4], Real *U)
{
    (void)s;
    culgt_dropin::read_ildg(file_name, SIZE, U, GF_PATTERN_GPU_TIMESLICE);
}

inline void writeILDG_timeslice(SiteCoord<4, FULL_SPLIT> s, const char *file_name, const char *output_name,
                                const short SIZE[4], Real *U, int steps)
{
    (void)s;
    culgt_dropin::write_ildg(file_name, output_name, SIZE, U, steps, GF_PATTERN_GPU_TIMESLICE);
}

#endif
