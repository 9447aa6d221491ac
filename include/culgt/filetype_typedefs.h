// culgt/filetype_typedefs.h -- lattice/filetypes/filetype_typedefs.h:26-97:
// the file type and real reinterpretation options of the apps, with the same
// case-insensitive stream operators (no boost needed).
#ifndef CULGT_FILETYPE_TYPEDEFS_H_
#define CULGT_FILETYPE_TYPEDEFS_H_

#include <cctype>
#include <istream>
#include <ostream>
#include <string>

#include "../su3gf_io.h"

enum ReinterpretReal { STANDARD = GF_REAL_STANDARD, DOUBLE = GF_REAL_DOUBLE, FLOAT = GF_REAL_FLOAT };
enum FileType { PLAIN = GF_FILE_PLAIN, HEADERONLY = GF_FILE_HEADERONLY, VOGT = GF_FILE_VOGT, ILDG = GF_FILE_ILDG,
                QCDSTAG = GF_FILE_QCDSTAG };

namespace culgt_dropin {
inline bool iequals(const std::string &a, const char *b)
{
    size_t n = 0;
    for (; b[n]; n++)
        if (n >= a.size() || std::tolower((unsigned char)a[n]) != std::tolower((unsigned char)b[n])) return false;
    return n == a.size();
}
}  // namespace culgt_dropin

inline std::istream &operator>>(std::istream &in, ReinterpretReal &t)
{
    std::string token;
    in >> token;
    if (culgt_dropin::iequals(token, "STANDARD")) t = STANDARD;
    else if (culgt_dropin::iequals(token, "DOUBLE")) t = DOUBLE;
    else if (culgt_dropin::iequals(token, "FLOAT")) t = FLOAT;
    return in;
}

inline std::ostream &operator<<(std::ostream &out, ReinterpretReal &t)
{
    return out << (t == STANDARD ? "STANDARD" : t == DOUBLE ? "DOUBLE" : t == FLOAT ? "FLOAT" : "");
}

inline std::istream &operator>>(std::istream &in, FileType &t)
{
    std::string token;
    in >> token;
    if (culgt_dropin::iequals(token, "PLAIN")) t = PLAIN;
    else if (culgt_dropin::iequals(token, "HEADERONLY")) t = HEADERONLY;
    else if (culgt_dropin::iequals(token, "VOGT")) t = VOGT;
    else if (culgt_dropin::iequals(token, "ILDG")) t = ILDG;
    else if (culgt_dropin::iequals(token, "QCDSTAG")) t = QCDSTAG;
    return in;
}

inline std::ostream &operator<<(std::ostream &out, FileType &t)
{
    return out << (t == PLAIN ? "PLAIN" : t == HEADERONLY ? "HEADERONLY" : t == VOGT ? "VOGT" : t == ILDG ? "ILDG"
                                                                                             : t == QCDSTAG ? "QCDSTAG" : "");
}

#endif
