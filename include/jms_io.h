/* jms_io.h -- on-disk formats of the JMSBench1 driver (test/JMSBench1):
 * EMS/SLI phase-grating files and the raw .bin STEM result channels, so that
 * the reference's inputs drive this library and its outputs read like the
 * reference's.  Host code.
 *
 * EMS file (prm_slice_file_info.h:37-55, 63-143): an 8192-byte header
 * (little-endian ints/floats at fixed offsets; a file whose first int is
 * outside [0, 16384] is byte-swapped, prm_slice_file_info.cpp:185-196) and
 * num_var complex64 grids at data_offset.  As the reference reads it, only the
 * header is byte-swapped, never the data (prm_slice.cpp:285-287).
 */
#ifndef JMS_IO_H
#define JMS_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jms_ems_header {
  uint32_t grid_x, grid_y;      /* grid size */
  uint32_t version;             /* extended header version (2010112301, 2012071101, 2020052801) */
  uint32_t num_var;             /* number of variants */
  uint32_t data_type;           /* content type: 0 phase grating, 1 potential */
  uint64_t data_offset;         /* byte offset of the data (default 8192) */
  uint64_t structure_offset;    /* byte offset of the atom table (default 4352) */
  float ekv;                    /* electron energy [keV] */
  float sx, sy, sz;             /* slice size x, y and thickness [nm] */
  char title[40];               /* slice name */
  int32_t byte_swapped;         /* 1 when the header was read byte-swapped */
} jms_ems_header;

/* read the header (prm_slice_file_info::load_ems_header, .cpp:79-149);
 * 0 = ok, >0 = error (4: grid edges outside [1, 16384] or no variant) */
int jms_ems_read_header(const char* path, jms_ems_header* h);
/* read the 2*grid_x*grid_y*num_var floats of the data block
 * (prm_slice::load_ems_data, prm_slice.cpp:253-300) */
int jms_ems_read_data(const char* path, const jms_ems_header* h, float* data);
/* write a phase-grating file: the header as save_ems_header writes it
 * (prm_slice_file_info.cpp:151-183, current version 2020052801), an empty
 * atom table at structure_offset, and the data at data_offset */
int jms_ems_write(const char* path, const jms_ems_header* h, const float* data);

/* STEM images from per-position integrated results res[pos][idetslc][idet]
 * (npos = nscan_x * nscan_y, scan index x fastest), divided by weight, in
 * the result layout of the bench (multislice.cpp:901-916):
 * out[pos + npos * (idetslc + ndetslc * idet)] */
int jms_stem_images(int npos, int ndetslc, int ndet, const float* res, float weight, float* out);
/* raw .bin of one detector channel of such images (prm_result::save,
 * prm_result.cpp:192-290): npos * ndetslc floats starting at channel idet */
int jms_bin_write_channel(const char* path, const float* images, int npos, int ndetslc, int ndet, int idet);

#ifdef __cplusplus
}
#endif
#endif /* JMS_IO_H */
