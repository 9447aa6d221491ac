// culgt/FileHeaderOnly.hxx -- lattice/filetypes/FileHeaderOnly.hxx: defined in culgt/FileTypes.hxx.
#include "FileTypes.hxx"
