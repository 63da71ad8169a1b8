// view_packing executable: the reference's bin-packing tool
// (src/tools/view_packing.cc:326-361) over libvsgm.so.
#include "vsgm.h"

int main(int argc, char** argv) { return vsgm_view_packing(argc, argv); }
