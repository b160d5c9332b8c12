/* Test infrastructure only.  Instantiates the reference's own vendored
 * stb_image_resize (include/third_party/stb_image_resize.h, v0.97) as a
 * shared library, so that the resize restatement in cgs_oracle.c can be
 * checked against the library the reference calls.  Built by `make ref`
 * from the header where it lies in the reference tree (nothing is copied);
 * the output goes to oracle/_ref/. */
#define STB_IMAGE_RESIZE_IMPLEMENTATION
#include "stb_image_resize.h"
