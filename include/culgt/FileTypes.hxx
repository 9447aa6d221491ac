// culgt/FileTypes.hxx -- the file types LinkFile is instantiated with:
// lattice/filetypes/FilePlain.hxx, FileHeaderOnly.hxx and FileVogt.hxx.  The
// reference's classes read and write their headers themselves; here they name
// the format, and the header a file type keeps between load and save lives in
// the library's file handle (include/su3gf_io.h).
#ifndef CULGT_FILETYPES_HXX_
#define CULGT_FILETYPES_HXX_

#include "../su3gf_io.h"

class FilePlain
{
public:
    static const int type = GF_FILE_PLAIN;
    explicit FilePlain(int /*LENGTH_OF_REAL*/ = 0) {}
};

class FileHeaderOnly
{
public:
    static const int type = GF_FILE_HEADERONLY;
    explicit FileHeaderOnly(int /*LENGTH_OF_REAL*/ = 0) {}
};

class FileVogt
{
public:
    static const int type = GF_FILE_VOGT;
    explicit FileVogt(int /*LENGTH_OF_REAL*/ = 0) {}
};

#endif
