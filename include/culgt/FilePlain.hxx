// culgt/FilePlain.hxx -- lattice/filetypes/FilePlain.hxx: defined in culgt/FileTypes.hxx.
#include "FileTypes.hxx"
