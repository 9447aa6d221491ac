// culgt/LinkFile.hxx -- drop-in for the reference's lattice/LinkFile.hxx:60-266:
// loads and saves a host configuration (all 18 reals per link, Real) through
// the library's file handle (include/su3gf_io.h).  FileType is FilePlain,
// FileHeaderOnly or FileVogt; FilePattern the reference's StandardPattern (the
// order of every configuration file the apps read); MemoryPattern GpuPattern
// or GpuPatternTimeslice; TheSite supplies the lattice extents (site.size).
// The header a file type keeps between load and save is kept by the handle.
// load / save return false on failure, as the reference's do, after printing
// the library's message.
#ifndef CULGT_LINKFILE_HXX_
#define CULGT_LINKFILE_HXX_

#include <cstdio>
#include <string>

#include "FileTypes.hxx"
#include "GlobalConstants.h"
#include "GpuPattern.hxx"
#include "filetype_typedefs.h"
#include "../su3gf_io.h"

namespace culgt_dropin {
template <class TheSite> gf_lattice file_lattice(const TheSite &site)
{
    gf_lattice l;
    for (int i = 0; i < 4; i++) l.dims[i] = site.size[i];
    l.dtype = sizeof(Real) == sizeof(double) ? GF_F64 : GF_F32;
    l.layout = GF_LAYOUT_PATTERN18;
    return l;
}

inline bool file_ok(gf_status st, const char *what, const std::string &filename)
{
    if (st == GF_OK) return true;
    fprintf(stderr, "%s %s: %s\n", what, filename.c_str(), gf_last_error());
    return false;
}

// a file handle for one call (the ILDG / QCDSTAG functions of the apps)
struct FileHandle {
    gf_linkfile *h = nullptr;
    FileHandle(int type, int pattern) { check(gf_linkfile_create(type, GF_REAL_STANDARD, pattern, &h)); }
    ~FileHandle() { gf_linkfile_destroy(h); }
    FileHandle(const FileHandle &) = delete;
    FileHandle &operator=(const FileHandle &) = delete;
};
}  // namespace culgt_dropin

template <class FileType, class FilePattern, class MemoryPattern, class TheSite> class LinkFile
{
    static_assert(FilePattern::is_standard, "configuration files are read in the reference's StandardPattern order");

public:
    explicit LinkFile(ReinterpretReal reinterpret = STANDARD) : filetype(0)
    {
        culgt_dropin::check(gf_linkfile_create(FileType::type, reinterpret, MemoryPattern::memory_pattern, &h));
    }
    ~LinkFile() { gf_linkfile_destroy(h); }
    LinkFile(const LinkFile &) = delete;
    LinkFile &operator=(const LinkFile &) = delete;

    bool load(TheSite site, std::string filename, Real *U)
    {
        const gf_lattice l = culgt_dropin::file_lattice(site);
        return culgt_dropin::file_ok(gf_linkfile_load(h, filename.c_str(), &l, U), "cannot load", filename);
    }

    bool save(TheSite site, std::string filename, Real *U)
    {
        const gf_lattice l = culgt_dropin::file_lattice(site);
        return culgt_dropin::file_ok(gf_linkfile_save(h, filename.c_str(), &l, U, 0), "cannot save", filename);
    }

    FileType filetype;

private:
    gf_linkfile *h = nullptr;
};

#endif
