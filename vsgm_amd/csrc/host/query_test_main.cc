// query_test executable: the reference's plan printer
// (src/tools/query_test.cc:12-65) over libvsgm.so.
#include "vsgm.h"

int main(int argc, char** argv) { return vsgm_query_test(argc, argv); }
