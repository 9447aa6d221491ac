/*
 * su3gf_io.h -- gauge configurations on disk, in the formats of the
 * reference's applications.  Host code only: it reads and writes host
 * arrays and never touches a GPU.
 *
 * Replaces:
 *   lattice/LinkFile.hxx:60-266 (LinkFile<FileType, StandardPattern,
 *       GpuPattern | GpuPatternTimeslice, SiteCoord>::load / save) with the
 *       file types of lattice/filetypes/{FilePlain,FileHeaderOnly,FileVogt}.hxx;
 *   gaugefixing/apps/ildg.cpp:26-451 (readILDG, writeILDG and their
 *       _timeslice forms) and the c-lime record layer it calls;
 *   gaugefixing/apps/qcdstag.cpp:17-227 (readQCDSTAG, writeQCDSTAG and their
 *       _timeslice forms).
 *
 * The host array U holds all 18 reals of every link, of type float or double
 * (lat->dtype; lat->layout must be GF_LAYOUT_PATTERN18), in one of the two
 * memory patterns of the reference's apps:
 *   GF_PATTERN_GPU            GpuPattern (GpuPattern.hxx:53-55):
 *                             U[idx + V (c + 2 (j + 3 (i + 3 mu)))], idx the
 *                             parity-split site index;
 *   GF_PATTERN_GPU_TIMESLICE  GpuPatternTimeslice (GpuPatternTimeslice.hxx:57-61):
 *                             timeslice t at offset 72 V3 t, inside it
 *                             GpuPattern over {1, Nx, Ny, Nz}.
 * mu = 0 is the time direction.  Lattice sites are (t, x, y, z) with
 * extents {Nt, Nx, Ny, Nz} = lat->dims.
 *
 * File layouts (the reference's behaviour, restated):
 *   PLAIN       V * 72 reals, site-major in the reference's standard order:
 *               real u = c + 2 (j + 3 (i + 3 (mu + 4 n))),
 *               n = ((t Nx + x) Ny + y) Nz + z (StandardPattern.hxx:53-70);
 *               native byte order.  The real type is lat->dtype (STANDARD),
 *               or double / float (GF_REAL_DOUBLE / GF_REAL_FLOAT, converted
 *               on the way in and out; LinkFile.hxx:125-150).
 *   HEADERONLY  an arbitrary header, then the PLAIN array; the header length
 *               is the file length minus the array size
 *               (FileHeaderOnly.hxx:60-80).  save writes the header of the
 *               last file loaded or read (none if there was none).
 *   VOGT        int16 ndim, int16 nc, int16 latsize[ndim], int16 length of a
 *               real, then the PLAIN array (FileVogt.hxx:52-102).  load
 *               requires ndim 4, nc 3, a volume equal to lat's and a real
 *               length equal to the one in use (the reference reads on
 *               regardless); save writes the header of the last file loaded,
 *               else {4, 3, Nt, Nx, Ny, Nz, length of a real}.
 *   ILDG        LIME records; the record "ildg-binary-data" holds V * 72
 *               big-endian doubles, element
 *               ((t Nz + z) Ny + y) Nx + x) * 72 + 18 m + 6 i + 2 j + c
 *               for file direction m = x, y, z, t, i.e. mu = (m + 1) % 4
 *               (ildg.cpp:63-97).  Always double, whatever the reinterpretation.
 *               save writes the records of the last file loaded or read, in
 *               their order and with their MB / ME flags, the binary record
 *               replaced by U and every other record rewritten as its text up
 *               to the first NUL, "xlf-info" with " SA steps <sa_steps>"
 *               appended (ildg.cpp:101-204); with no such file it writes a
 *               minimal message: "ildg-format" (an ILDG XML description of
 *               the field) and "ildg-binary-data".
 *   QCDSTAG     V * 72 native doubles, direction-major:
 *               ((mu Nt + t) Nz + z) Ny + y) Nx + x) * 18 + 6 i + 2 j + c
 *               with the lattice's own mu (qcdstag.cpp:17-118).  Always double.
 *
 * Errors: GF_ERR_IO for a file that cannot be opened, is too short or is not
 * in the expected format; gf_last_error() holds the reason.
 */
#ifndef SU3GF_IO_H_
#define SU3GF_IO_H_

#include "su3gf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* filetype_typedefs.h:60 (FileType) */
typedef enum {
    GF_FILE_PLAIN = 0,
    GF_FILE_HEADERONLY = 1,
    GF_FILE_VOGT = 2,
    GF_FILE_ILDG = 3,
    GF_FILE_QCDSTAG = 4
} gf_filetype;

/* filetype_typedefs.h:29 (ReinterpretReal) */
typedef enum {
    GF_REAL_STANDARD = 0, /* the real type of the host array */
    GF_REAL_DOUBLE = 1,
    GF_REAL_FLOAT = 2
} gf_reinterpret;

typedef enum {
    GF_PATTERN_GPU = 0,
    GF_PATTERN_GPU_TIMESLICE = 1
} gf_memory_pattern;

/* A file handle of one type: it keeps the header (HEADERONLY, VOGT) or the
 * records (ILDG) of the last file loaded, for save, as a LinkFile object
 * keeps its file type's state. */
typedef struct gf_linkfile gf_linkfile;

gf_status gf_linkfile_create(int filetype, int reinterpret, int pattern, gf_linkfile **out);
void gf_linkfile_destroy(gf_linkfile *f);

/* LinkFile::load (LinkFile.hxx:84-165), readILDG, readQCDSTAG */
gf_status gf_linkfile_load(gf_linkfile *f, const char *path, const gf_lattice *lat, void *U);

/* LinkFile::save (LinkFile.hxx:171-255), writeILDG, writeQCDSTAG; sa_steps
 * only enters the ILDG "xlf-info" record */
gf_status gf_linkfile_save(gf_linkfile *f, const char *path, const gf_lattice *lat, const void *U, int sa_steps);

/* Keep the header (HEADERONLY, VOGT) or records (ILDG) of a file without
 * reading its links: writeILDG's template file (ildg.cpp:101), or the header
 * for a configuration that was generated rather than loaded. */
gf_status gf_linkfile_read_header(gf_linkfile *f, const char *path, const gf_lattice *lat);

#ifdef __cplusplus
}
#endif

#endif /* SU3GF_IO_H_ */
