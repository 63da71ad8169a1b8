// preprocess_graph executable: the reference's SNAP -> .bin tool
// (src/preprocess/preprocess_graph.cc:9-36) over libvsgm.so.
#include "vsgm.h"

int main(int argc, char** argv) { return vsgm_preprocess_graph(argc, argv); }
