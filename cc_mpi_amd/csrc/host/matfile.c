/*
 * MAT-file level 5 sparse-matrix reader and writer.
 *
 * Written from MathWorks' public "MAT-File Format" document (level 5 MAT
 * files): a 128-byte header (text, subsystem offset, version 0x0100, endian
 * indicator "IM"/"MI"), then a sequence of data elements, each an 8-byte
 * tag (type, byte count) or a 4-byte "small element" tag (byte count in the
 * upper half-word) followed by data padded to 8 bytes.  A variable is an
 * miMATRIX element holding array flags, dimensions, name and the class's
 * data; miCOMPRESSED wraps one such element in a zlib stream.  Sparse
 * arrays are compressed sparse columns: ir (row index per stored entry),
 * jc (column starts, ncols + 1), pr (values; ignored here).
 *
 * Reading follows the reference's loader (src/graph.c:38-128): variables are
 * visited in file order; a struct named "Problem" yields its field "A"; else
 * the first sparse variable is taken.  Rows become vertices (n = dims[0]);
 * row r of the CSR lists, in increasing order, every column c whose CSC
 * column holds an entry at row r (rows out of [0, n) are dropped); values
 * are ignored.  Return codes follow the reference: 1 the file cannot be
 * opened (or is not a level 5 MAT file), 2 no sparse matrix found, 3 out
 * of memory.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>
#include <zlib.h>

#include "../../../include/graph.h"

enum
{
  miINT8 = 1,
  miUINT8 = 2,
  miINT16 = 3,
  miUINT16 = 4,
  miINT32 = 5,
  miUINT32 = 6,
  miSINGLE = 7,
  miDOUBLE = 9,
  miINT64 = 12,
  miUINT64 = 13,
  miMATRIX = 14,
  miCOMPRESSED = 15,
  miUTF8 = 16,
};

enum
{
  mxSTRUCT_CLASS = 2,
  mxCHAR_CLASS = 4,
  mxSPARSE_CLASS = 5,
};

#define MAT_FLAG_LOGICAL 0x0200u

/* ---- reading ------------------------------------------------------------ */

typedef struct
{
  const uint8_t *p; /* element data */
  uint64_t len;     /* bytes in the region */
  int swap;         /* file byte order differs from the host's */
} Region;

typedef struct
{
  uint32_t type;
  uint32_t nbytes;
  const uint8_t *data;
  uint64_t next; /* offset of the following element in the region */
} Elem;

static uint16_t rd16(const uint8_t *p, int swap)
{
  uint16_t v;
  memcpy(&v, p, 2);
  return swap ? (uint16_t)((v >> 8) | (v << 8)) : v;
}

static uint32_t rd32(const uint8_t *p, int swap)
{
  uint32_t v;
  memcpy(&v, p, 4);
  return swap ? __builtin_bswap32(v) : v;
}

static uint64_t rd64(const uint8_t *p, int swap)
{
  uint64_t v;
  memcpy(&v, p, 8);
  return swap ? __builtin_bswap64(v) : v;
}

static uint64_t pad8(uint64_t x) { return (x + 7u) & ~(uint64_t)7u; }

/* Element starting at `off` of region r; 0 on success. */
static int elem_at(const Region *r, uint64_t off, Elem *e)
{
  if (off + 8 > r->len)
    return -1;
  const uint32_t t0 = rd32(r->p + off, r->swap);
  if (t0 >> 16)
  { /* small data element: 2-byte count, 2-byte type, up to 4 data bytes */
    e->type = t0 & 0xFFFFu;
    e->nbytes = t0 >> 16;
    if (e->nbytes > 4)
      return -1;
    e->data = r->p + off + 4;
    e->next = off + 8;
    return 0;
  }
  e->type = t0;
  e->nbytes = rd32(r->p + off + 4, r->swap);
  e->data = r->p + off + 8;
  if (off + 8 + (uint64_t)e->nbytes > r->len)
    return -1;
  /* compressed elements are not padded; every other element is */
  e->next = off + 8 + (e->type == miCOMPRESSED ? (uint64_t)e->nbytes : pad8(e->nbytes));
  if (e->next > r->len)
    e->next = r->len;
  return 0;
}

static int type_size(uint32_t t)
{
  switch (t)
  {
  case miINT8: case miUINT8: case miUTF8: return 1;
  case miINT16: case miUINT16: return 2;
  case miINT32: case miUINT32: case miSINGLE: return 4;
  case miDOUBLE: case miINT64: case miUINT64: return 8;
  default: return 0;
  }
}

/* i-th entry of an integer-typed element as int64 (-1 for a non-integer type). */
static int64_t int_at(const Elem *e, uint64_t i, int swap)
{
  const uint8_t *q = e->data + i * (uint64_t)type_size(e->type);
  switch (e->type)
  {
  case miINT8: return (int8_t)q[0];
  case miUINT8: return q[0];
  case miINT16: return (int16_t)rd16(q, swap);
  case miUINT16: return rd16(q, swap);
  case miINT32: return (int32_t)rd32(q, swap);
  case miUINT32: return rd32(q, swap);
  case miINT64: return (int64_t)rd64(q, swap);
  case miUINT64: return (int64_t)rd64(q, swap);
  default: return -1;
  }
}

/* Header of one miMATRIX body (region = the element's data bytes). */
typedef struct
{
  uint32_t cls;
  uint32_t flags;
  int ndims;
  int64_t dims[8];
  char name[64];
  uint64_t body; /* offset of the first class-specific subelement */
} MatrixHead;

static int matrix_head(const Region *r, MatrixHead *h)
{
  Elem e;
  memset(h, 0, sizeof(*h));
  if (r->len == 0)
    return -1; /* empty matrix element (e.g. an unset struct field) */
  if (elem_at(r, 0, &e) || e.type != miUINT32 || e.nbytes < 8)
    return -1;
  h->flags = rd32(e.data, r->swap);
  h->cls = h->flags & 0xFFu;
  uint64_t off = e.next;
  if (elem_at(r, off, &e) || type_size(e.type) == 0)
    return -1;
  h->ndims = (int)(e.nbytes / (uint32_t)type_size(e.type));
  if (h->ndims < 2 || h->ndims > 8)
    return -1;
  for (int d = 0; d < h->ndims; ++d)
    h->dims[d] = int_at(&e, (uint64_t)d, r->swap);
  off = e.next;
  if (elem_at(r, off, &e) || (e.type != miINT8 && e.type != miUINT8))
    return -1;
  const uint32_t nl = e.nbytes < sizeof(h->name) - 1 ? e.nbytes : (uint32_t)sizeof(h->name) - 1;
  memcpy(h->name, e.data, nl);
  h->name[nl] = '\0';
  h->body = e.next;
  return 0;
}

/* Sparse arrays: locate ir and jc. */
static int sparse_parts(const Region *r, const MatrixHead *h, Elem *ir, Elem *jc)
{
  if (elem_at(r, h->body, ir) || elem_at(r, ir->next, jc))
    return -1;
  if (type_size(ir->type) == 0 || type_size(jc->type) == 0 || ir->type == miSINGLE ||
      ir->type == miDOUBLE || jc->type == miSINGLE || jc->type == miDOUBLE)
    return -1;
  return 0;
}

/* Struct arrays: the miMATRIX region of field `want` of element 0. */
static int struct_field(const Region *r, const MatrixHead *h, const char *want, Region *out)
{
  Elem e;
  if (elem_at(r, h->body, &e) || type_size(e.type) == 0 || e.nbytes == 0)
    return -1;
  const int64_t flen = int_at(&e, 0, r->swap);
  uint64_t off = e.next;
  if (flen <= 0 || elem_at(r, off, &e) || (e.type != miINT8 && e.type != miUINT8))
    return -1;
  const uint64_t nfields = e.nbytes / (uint64_t)flen;
  const uint8_t *names = e.data;
  off = e.next;
  for (uint64_t f = 0; f < nfields; ++f)
  {
    Elem fe;
    if (elem_at(r, off, &fe) || fe.type != miMATRIX)
      return -1;
    const char *nm = (const char *)names + f * (uint64_t)flen;
    if (strnlen(nm, (size_t)flen) == strlen(want) && strncmp(nm, want, strlen(want)) == 0)
    {
      out->p = fe.data;
      out->len = fe.nbytes;
      out->swap = r->swap;
      return 0;
    }
    off = fe.next;
  }
  return -1;
}

/* CSC (ir, jc) of an n x ncols matrix -> CSR with rows = matrix rows. */
static int csc_to_csr(const Elem *ir, const Elem *jc, int swap, uint32_t n, uint32_t ncols,
                      CSRGraph *out)
{
  const uint64_t njc = jc->nbytes / (uint64_t)type_size(jc->type);
  const uint64_t nir = ir->nbytes / (uint64_t)type_size(ir->type);
  if (njc < (uint64_t)ncols + 1)
    return 2;
  const int64_t nz = int_at(jc, ncols, swap);
  if (nz < 0 || (uint64_t)nz > nir)
    return 2;
  uint64_t *row_ptr = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
  uint32_t *col_idx = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(nz ? nz : 1));
  uint64_t *head = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
  if (!row_ptr || !col_idx || !head)
  {
    free(row_ptr);
    free(col_idx);
    free(head);
    return 3;
  }
  /* fast path: int32 row indices in host byte order, read in place */
  const int32_t *ir32 = (ir->type == miINT32 && !swap) ? (const int32_t *)ir->data : NULL;
  for (int pass = 0; pass < 2; ++pass)
  {
    int64_t j0 = int_at(jc, 0, swap);
    for (uint32_t c = 0; c < ncols; ++c)
    {
      const int64_t j1 = int_at(jc, (uint64_t)c + 1, swap);
      if (j0 < 0 || j1 < j0 || j1 > nz)
      {
        free(row_ptr);
        free(col_idx);
        free(head);
        return 2;
      }
      for (int64_t j = j0; j < j1; ++j)
      {
        int32_t rr;
        if (ir32)
          memcpy(&rr, ir32 + j, sizeof(rr));
        else
          rr = (int32_t)int_at(ir, (uint64_t)j, swap);
        if (rr < 0 || (uint32_t)rr >= n)
          continue;
        if (pass == 0)
          row_ptr[rr + 1]++;
        else
          col_idx[head[rr]++] = c;
      }
      j0 = j1;
    }
    if (pass == 0)
    {
      for (uint32_t i = 0; i < n; ++i)
        row_ptr[i + 1] += row_ptr[i];
      memcpy(head, row_ptr, sizeof(uint64_t) * (size_t)n);
    }
  }
  free(head);
  out->n = n;
  out->m = row_ptr[n];
  out->row_ptr = row_ptr;
  out->col_idx = col_idx;
  return 0;
}

/* Take a variable if it is the one the loader wants: 0 taken (out filled),
 * 1 not wanted, 2/3 error codes of the loader. */
static int try_variable(const Region *r, CSRGraph *out)
{
  MatrixHead h;
  if (matrix_head(r, &h))
    return 1;
  Region body = *r;
  if (h.cls == mxSTRUCT_CLASS && strcmp(h.name, "Problem") == 0)
  {
    /* the reference stops at a "Problem" struct whether or not it has a
     * sparse A; without one there is no matrix to read */
    Region fa;
    if (struct_field(r, &h, "A", &fa) || matrix_head(&fa, &h) || h.cls != mxSPARSE_CLASS)
      return 2;
    body = fa;
  }
  else if (h.cls != mxSPARSE_CLASS)
  {
    return 1;
  }
  if (h.ndims != 2 || h.dims[0] < 0 || h.dims[1] < 0 || h.dims[0] > UINT32_MAX ||
      h.dims[1] > UINT32_MAX)
    return 2;
  Elem ir, jc;
  if (sparse_parts(&body, &h, &ir, &jc))
    return 2;
  return csc_to_csr(&ir, &jc, body.swap, (uint32_t)h.dims[0], (uint32_t)h.dims[1], out);
}

/* Inflate a miCOMPRESSED payload completely; returns malloc'd bytes. */
static uint8_t *inflate_all(const uint8_t *src, uint64_t n, int swap, uint64_t *out_len)
{
  z_stream z;
  memset(&z, 0, sizeof(z));
  if (inflateInit(&z) != Z_OK)
    return NULL;
  uint8_t tag[8];
  z.next_in = (Bytef *)src;
  z.avail_in = (uInt)n;
  z.next_out = tag;
  z.avail_out = 8;
  int rc = inflate(&z, Z_SYNC_FLUSH);
  if (z.avail_out != 0 || (rc != Z_OK && rc != Z_STREAM_END))
  {
    inflateEnd(&z);
    return NULL;
  }
  const uint64_t inner = rd32(tag + 4, swap);
  const uint64_t total = 8 + pad8(inner);
  uint8_t *buf = (uint8_t *)malloc((size_t)total);
  if (!buf)
  {
    inflateEnd(&z);
    return NULL;
  }
  memcpy(buf, tag, 8);
  uint64_t have = 8;
  while (have < total && rc != Z_STREAM_END)
  {
    const uint64_t room = total - have;
    z.next_out = buf + have;
    z.avail_out = (uInt)(room > (1u << 30) ? (1u << 30) : room);
    const uInt before = z.avail_out;
    rc = inflate(&z, Z_NO_FLUSH);
    have += before - z.avail_out;
    if (rc != Z_OK && rc != Z_STREAM_END)
    {
      free(buf);
      inflateEnd(&z);
      return NULL;
    }
    if (rc == Z_OK && before == z.avail_out && z.avail_in == 0)
      break; /* truncated stream */
  }
  inflateEnd(&z);
  if (have < 8 + inner)
  {
    free(buf);
    return NULL;
  }
  *out_len = have;
  return buf;
}

int load_csr_from_mat(const char *path, CSRGraph *out)
{
  memset(out, 0, sizeof(*out));
  const int fd = open(path, O_RDONLY);
  if (fd < 0)
  {
    fprintf(stderr, "Error opening .mat file: %s\n", path);
    return 1;
  }
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size < 128)
  {
    close(fd);
    fprintf(stderr, "Error opening .mat file: %s\n", path);
    return 1;
  }
  const uint64_t size = (uint64_t)st.st_size;
  uint8_t *map = (uint8_t *)mmap(NULL, (size_t)size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (map == MAP_FAILED)
  {
    fprintf(stderr, "Error opening .mat file: %s\n", path);
    return 1;
  }
  /* endian indicator: the characters 'M','I' written as one 16-bit value */
  uint16_t ind;
  memcpy(&ind, map + 126, 2);
  int swap;
  if (ind == (('M' << 8) | 'I'))
    swap = 0;
  else if (ind == (('I' << 8) | 'M'))
    swap = 1;
  else
  {
    munmap(map, (size_t)size);
    fprintf(stderr, "Error opening .mat file: %s\n", path);
    return 1;
  }
  if (rd16(map + 124, swap) != 0x0100)
  {
    munmap(map, (size_t)size);
    fprintf(stderr, "Error opening .mat file: %s\n", path);
    return 1;
  }
  const Region file = {map, size, swap};
  int rc = 2;
  for (uint64_t off = 128; off + 8 <= size;)
  {
    Elem e;
    if (elem_at(&file, off, &e))
      break;
    if (e.type == miMATRIX)
    {
      const Region r = {e.data, e.nbytes, swap};
      const int t = try_variable(&r, out);
      if (t != 1)
      {
        rc = t;
        break;
      }
    }
    else if (e.type == miCOMPRESSED)
    {
      uint64_t len = 0;
      uint8_t *buf = inflate_all(e.data, e.nbytes, swap, &len);
      if (!buf)
        break; /* unreadable variable: stop as a failed read would */
      const Region whole = {buf, len, swap};
      Elem in;
      int t = 1;
      if (elem_at(&whole, 0, &in) == 0 && in.type == miMATRIX)
      {
        const Region r = {in.data, in.nbytes, swap};
        t = try_variable(&r, out);
      }
      free(buf);
      if (t != 1)
      {
        rc = t;
        break;
      }
    }
    off = e.next;
  }
  munmap(map, (size_t)size);
  if (rc == 2)
    fprintf(stderr, "No sparse matrix found in .mat file %s\n", path);
  else if (rc == 3)
    fprintf(stderr, "Memory allocation failed\n");
  return rc;
}

int load_csr_from_file(const char *path, int symmetrize, int drop_self_loops, CSRGraph *out)
{
  (void)symmetrize;
  (void)drop_self_loops;
  const char *ext = strrchr(path, '.');
  if (!ext)
    ext = "";
  if (strcasecmp(ext, ".mtx") == 0 || strcasecmp(ext, ".txt") == 0)
    return -1;
  if (strcasecmp(ext, ".mat") == 0)
    return load_csr_from_mat(path, out);
  fprintf(stderr, "Unsupported file extension: %s\n", ext);
  return 99;
}

/* ---- writing ------------------------------------------------------------ */

typedef struct
{
  uint8_t *p;
  uint64_t len, cap;
  int swap;
  int oom;
} Out;

static void put(Out *o, const void *src, uint64_t n)
{
  if (o->oom)
    return;
  if (o->len + n > o->cap)
  {
    uint64_t cap = o->cap ? o->cap : 4096;
    while (cap < o->len + n)
      cap *= 2;
    uint8_t *q = (uint8_t *)realloc(o->p, (size_t)cap);
    if (!q)
    {
      o->oom = 1;
      return;
    }
    o->p = q;
    o->cap = cap;
  }
  memcpy(o->p + o->len, src, (size_t)n);
  o->len += n;
}

static void put32(Out *o, uint32_t v)
{
  if (o->swap)
    v = __builtin_bswap32(v);
  put(o, &v, 4);
}

static void put16(Out *o, uint16_t v)
{
  if (o->swap)
    v = (uint16_t)((v >> 8) | (v << 8));
  put(o, &v, 2);
}

static void put_pad(Out *o)
{
  static const uint8_t zero[8] = {0};
  put(o, zero, pad8(o->len) - o->len);
}

static void put_tag(Out *o, uint32_t type, uint32_t nbytes)
{
  put32(o, type);
  put32(o, nbytes);
}

/* Name subelement: small-element form when it fits in 4 bytes. */
static void put_name(Out *o, const char *name)
{
  const uint32_t n = (uint32_t)strlen(name);
  if (n > 0 && n <= 4)
  {
    put32(o, (n << 16) | miINT8);
    uint8_t b[4] = {0};
    memcpy(b, name, n);
    put(o, b, 4);
    return;
  }
  put_tag(o, miINT8, n);
  put(o, name, n);
  put_pad(o);
}

/* Backpatch a 32-bit value at byte offset `at`. */
static void patch32(Out *o, uint64_t at, uint32_t v)
{
  if (o->oom)
    return;
  if (o->swap)
    v = __builtin_bswap32(v);
  memcpy(o->p + at, &v, 4);
}

/* A 1 x 1 sparse logical n x n matrix whose row r holds g's row r. */
static int put_sparse(Out *o, const char *name, const CSRGraph *g)
{
  const uint32_t n = g->n;
  const uint64_t nz = g->m;
  if (nz > (uint64_t)INT32_MAX || (nz * 4 + 8) > UINT32_MAX)
    return 4; /* level 5 stores jc as int32 and element sizes in 32 bits */
  for (uint64_t k = 0; k < nz; ++k)
    if (g->col_idx[k] >= n)
      return 4;
  /* CSC of the matrix = transpose of the CSR: column c lists rows r with
   * c in row r, in increasing r */
  uint64_t *jc = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
  int32_t *ir = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nz ? nz : 1));
  if (!jc || !ir)
  {
    free(jc);
    free(ir);
    return 3;
  }
  for (uint64_t k = 0; k < nz; ++k)
    jc[g->col_idx[k] + 1]++;
  for (uint32_t c = 0; c < n; ++c)
    jc[c + 1] += jc[c];
  uint64_t *head = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
  if (!head)
  {
    free(jc);
    free(ir);
    return 3;
  }
  memcpy(head, jc, sizeof(uint64_t) * (size_t)n);
  for (uint32_t r = 0; r < n; ++r)
    for (uint64_t k = g->row_ptr[r]; k < g->row_ptr[r + 1]; ++k)
      ir[head[g->col_idx[k]]++] = (int32_t)r;
  free(head);

  put_tag(o, miMATRIX, 0);
  const uint64_t size_at = o->len - 4, start = o->len;
  put_tag(o, miUINT32, 8);
  put32(o, mxSPARSE_CLASS | MAT_FLAG_LOGICAL);
  put32(o, (uint32_t)(nz ? nz : 1)); /* nzmax */
  put_tag(o, miINT32, 8);
  put32(o, n);
  put32(o, n);
  put_name(o, name);
  put_tag(o, miINT32, (uint32_t)(nz * 4));
  for (uint64_t k = 0; k < nz; ++k)
    put32(o, (uint32_t)ir[k]);
  put_pad(o);
  put_tag(o, miINT32, (uint32_t)(((uint64_t)n + 1) * 4));
  for (uint32_t c = 0; c <= n; ++c)
    put32(o, (uint32_t)jc[c]);
  put_pad(o);
  put_tag(o, miUINT8, (uint32_t)nz); /* logical values: all ones */
  for (uint64_t k = 0; k < nz; ++k)
  {
    const uint8_t one = 1;
    put(o, &one, 1);
  }
  put_pad(o);
  patch32(o, size_at, (uint32_t)(o->len - start));
  free(jc);
  free(ir);
  return o->oom ? 3 : 0;
}

/* A 1 x k char array (UTF-16 code units, as MATLAB writes them). */
static void put_char(Out *o, const char *name, const char *text)
{
  const uint32_t k = (uint32_t)strlen(text);
  put_tag(o, miMATRIX, 0);
  const uint64_t size_at = o->len - 4, start = o->len;
  put_tag(o, miUINT32, 8);
  put32(o, mxCHAR_CLASS);
  put32(o, 0);
  put_tag(o, miINT32, 8);
  put32(o, 1);
  put32(o, k);
  put_name(o, name);
  put_tag(o, miUINT16, k * 2);
  for (uint32_t i = 0; i < k; ++i)
    put16(o, (uint16_t)(uint8_t)text[i]);
  put_pad(o);
  patch32(o, size_at, (uint32_t)(o->len - start));
}

int cc_mat_write_sparse(const char *path, const char *name, const CSRGraph *g, int flags)
{
  if (!path || !name || !g || (g->m && (!g->row_ptr || !g->col_idx)) || !g->row_ptr)
    return 4;
  Out o = {NULL, 0, 0, 0, 0};
  const uint16_t probe = 1;
  const int host_le = *(const uint8_t *)&probe == 1;
  const int want_be = (flags & CC_MAT_BIG_ENDIAN) != 0;
  o.swap = want_be ? host_le : !host_le;

  /* header: 116 text bytes, 8 subsystem bytes, version, endian indicator */
  char text[117];
  memset(text, ' ', sizeof(text));
  time_t now = time(NULL);
  struct tm tmv;
  gmtime_r(&now, &tmv);
  char when[64];
  strftime(when, sizeof(when), "%a %b %d %H:%M:%S %Y", &tmv);
  const int tl = snprintf(text, sizeof(text), "MATLAB 5.0 MAT-file, Platform: cc_mpi_amd, Created on: %s", when);
  if (tl >= 0 && tl < 116)
    text[tl] = ' ';
  put(&o, text, 116);
  const uint8_t sub[8] = {0};
  put(&o, sub, 8);
  put16(&o, 0x0100);
  put16(&o, ('M' << 8) | 'I');

  Out var = {NULL, 0, 0, o.swap, 0};
  int rc;
  if (flags & CC_MAT_PROBLEM)
  {
    /* struct Problem with fields "name" (char) and "A" (the sparse matrix) */
    put_tag(&var, miMATRIX, 0);
    const uint64_t size_at = var.len - 4, start = var.len;
    put_tag(&var, miUINT32, 8);
    put32(&var, mxSTRUCT_CLASS);
    put32(&var, 0);
    put_tag(&var, miINT32, 8);
    put32(&var, 1);
    put32(&var, 1);
    put_name(&var, "Problem");
    const uint32_t flen = 32;
    put32(&var, (4u << 16) | miINT32);
    put32(&var, flen);
    put_tag(&var, miINT8, 2 * flen);
    char fields[64] = {0};
    memcpy(fields, "name", 4);
    memcpy(fields + flen, "A", 1);
    put(&var, fields, 2 * flen);
    put_char(&var, "", name);
    rc = put_sparse(&var, "", g);
    patch32(&var, size_at, (uint32_t)(var.len - start));
  }
  else
  {
    rc = put_sparse(&var, name, g);
  }
  if (rc == 0 && var.oom)
    rc = 3;
  if (rc == 0)
  {
    if (flags & CC_MAT_COMPRESS)
    {
      uLongf zl = compressBound((uLong)var.len);
      uint8_t *z = (uint8_t *)malloc((size_t)zl);
      if (!z || compress2(z, &zl, var.p, (uLong)var.len, 6) != Z_OK || zl > UINT32_MAX)
        rc = 3;
      else
      {
        put_tag(&o, miCOMPRESSED, (uint32_t)zl);
        put(&o, z, zl);
      }
      free(z);
    }
    else
    {
      put(&o, var.p, var.len);
    }
  }
  free(var.p);
  if (rc == 0 && o.oom)
    rc = 3;
  if (rc == 0)
  {
    FILE *f = fopen(path, "wb");
    if (!f)
      rc = 1;
    else
    {
      if (fwrite(o.p, 1, (size_t)o.len, f) != o.len)
        rc = 1;
      if (fclose(f) != 0)
        rc = 1;
    }
  }
  free(o.p);
  return rc;
}
