/* Dumps one MAS HDF5 file (the reference's input format: float datasets
 * dim1 = r [Rs], dim2 = theta, dim3 = phi and Data [phi][theta][r],
 * src/masIO.c:34-240) as raw little-endian records:
 *   int32 n1, n2, n3, then float dim1[n1], dim2[n2], dim3[n3], Data[n3*n2*n1].
 * Built and run by make_fixtures.py against the HDF5 C library in /opt/conda;
 * used only to write test fixtures. */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>

static float *read_all(hid_t f, const char *name, hsize_t *count) {
  hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
  if (d < 0) return NULL;
  hid_t s = H5Dget_space(d);
  hssize_t n = H5Sget_simple_extent_npoints(s);
  float *buf = malloc((size_t)n * sizeof(float));
  if (buf && H5Dread(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) {
    free(buf);
    buf = NULL;
  }
  H5Sclose(s);
  H5Dclose(d);
  *count = (hsize_t)n;
  return buf;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: mas_h5dump in.h5 out.bin\n");
    return 2;
  }
  hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return 1;
  const char *names[4] = {"dim1", "dim2", "dim3", "Data"};
  float *arr[4];
  hsize_t n[4];
  for (int i = 0; i < 4; i++) {
    arr[i] = read_all(f, names[i], &n[i]);
    if (!arr[i]) return 1;
  }
  H5Fclose(f);
  if (n[3] != n[0] * n[1] * n[2]) return 1;
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 1;
  for (int i = 0; i < 3; i++) {
    int v = (int)n[i];
    fwrite(&v, 4, 1, o);
  }
  for (int i = 0; i < 4; i++) fwrite(arr[i], 4, n[i], o);
  fclose(o);
  return 0;
}
