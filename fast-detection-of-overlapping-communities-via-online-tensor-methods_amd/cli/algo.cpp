// Placeholder CLI; replaced by the ./algo GPU_A driver once the pipeline entry points exist.
#include <cstdio>
#include "mmsb_stpm.h"
int main() { std::printf("%s\n", mmsb_version()); return 0; }
