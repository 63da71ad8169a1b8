// pm_pipeline executable: the reference's CLI entry point
// (src/app/pattern_matching_pipeline.cu:27) over libvsgm.so.
#include "vsgm.h"

int main(int argc, char** argv) {
  if (argc == 1) return -1;
  return vsgm_pm_pipeline(argc, argv);
}
