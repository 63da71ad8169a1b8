// kmeans executable: the reference's clustering tool
// (src/tools/kmeans.cc:636-694) over libvsgm.so, on the GPU.
#include "vsgm.h"

int main(int argc, char** argv) { return vsgm_kmeans_tool(argc, argv); }
