// culgt/FileVogt.hxx -- lattice/filetypes/FileVogt.hxx: defined in culgt/FileTypes.hxx.
#include "FileTypes.hxx"
