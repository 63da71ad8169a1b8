// fennel executable: the reference's streaming partitioner
// (src/tools/fennel.cc:21-101) over libvsgm.so.
#include "vsgm.h"

int main(int argc, char** argv) { return vsgm_fennel_tool(argc, argv); }
