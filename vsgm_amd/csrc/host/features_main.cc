// features executable: the reference's 1-hop feature graph tool
// (src/tools/features.cc:68-207) over libvsgm.so.
#include "vsgm.h"

int main(int argc, char** argv) { return vsgm_features_tool(argc, argv); }
