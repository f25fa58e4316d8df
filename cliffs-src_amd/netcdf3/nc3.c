/*
 * nc3 -- NetCDF-3 classic file reader/writer (see nc3.h).
 *
 * Format: the published NetCDF classic format specification (CDF-1, and
 * CDF-2 with 64-bit variable offsets): a big-endian header of dimension,
 * attribute and variable lists, then the fixed-size variables in definition
 * order, then the records, each holding one slab of every record variable.
 * Every header item and every fixed-size variable is padded to 4 bytes.
 */
#include "nc3.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXFILES 64
#define MAXDIMS 1024
#define MAXVARDIMS 32
#define MAXNAME 256
#define BADMAGIC (-1000) /* internal: not a classic file at all (no retry) */

typedef struct {
    char name[MAXNAME];
    int type;
    size_t n;
    unsigned char *data; /* host byte order */
} att_t;

typedef struct {
    int n, cap;
    att_t *a;
} atts_t;

typedef struct {
    char name[MAXNAME];
    size_t len; /* 0: the unlimited dimension */
} dim_t;

typedef struct {
    char name[MAXNAME];
    int ndims;
    int dimids[MAXVARDIMS];
    int type;
    atts_t atts;
    uint64_t begin;
    uint64_t vsize; /* bytes of one variable (fixed) or one record slab, padded */
    int is_rec;
} var_t;

typedef struct {
    int used;
    FILE *fp;
    int writable, define, version;
    size_t numrecs;
    int ndims, nvars;
    dim_t *dims;
    var_t *vars;
    atts_t gatts;
    int unlim;
    uint64_t recsize, recbegin;
} file_t;

static file_t files[MAXFILES];

/* ---- small helpers ---------------------------------------------------- */

static size_t tsize(int t) {
    switch (t) {
        case NC3_BYTE:
        case NC3_CHAR: return 1;
        case NC3_SHORT: return 2;
        case NC3_INT:
        case NC3_FLOAT: return 4;
        case NC3_DOUBLE: return 8;
    }
    return 0;
}

static size_t pad4(size_t n) { return (n + 3) & ~(size_t)3; }

static int host_big(void) {
    const uint16_t x = 1;
    return *(const unsigned char *)&x == 0;
}

/* big-endian <-> host for n items of size s, in place */
static void swap_items(unsigned char *p, size_t n, size_t s) {
    if (s == 1 || host_big()) return;
    for (size_t i = 0; i < n; i++, p += s)
        for (size_t k = 0; k < s / 2; k++) {
            const unsigned char t = p[k];
            p[k] = p[s - 1 - k];
            p[s - 1 - k] = t;
        }
}

static file_t *get(int ncid) {
    if (ncid < 0 || ncid >= MAXFILES || !files[ncid].used) return NULL;
    return &files[ncid];
}

static void free_atts(atts_t *a) {
    for (int i = 0; i < a->n; i++) free(a->a[i].data);
    free(a->a);
    a->a = NULL;
    a->n = a->cap = 0;
}

static void free_file(file_t *f) {
    if (f->fp) fclose(f->fp);
    for (int v = 0; v < f->nvars; v++) free_atts(&f->vars[v].atts);
    free(f->vars);
    free(f->dims);
    free_atts(&f->gatts);
    memset(f, 0, sizeof(*f));
}

static int new_file(int *ncid) {
    for (int i = 0; i < MAXFILES; i++)
        if (!files[i].used) {
            memset(&files[i], 0, sizeof(files[i]));
            files[i].used = 1;
            files[i].unlim = -1;
            *ncid = i;
            return NC3_NOERR;
        }
    return NC3_ENFILE;
}

static atts_t *atts_of(file_t *f, int varid) {
    if (varid == NC3_GLOBAL) return &f->gatts;
    if (varid < 0 || varid >= f->nvars) return NULL;
    return &f->vars[varid].atts;
}

static int find_att(const atts_t *a, const char *name) {
    for (int i = 0; i < a->n; i++)
        if (strcmp(a->a[i].name, name) == 0) return i;
    return -1;
}

static int add_att(atts_t *a, const char *name, int type, size_t n, const void *values) {
    if (strlen(name) >= MAXNAME) return NC3_EMAXNAME;
    if (!tsize(type)) return NC3_EBADTYPE;
    int k = find_att(a, name);
    if (k < 0) {
        if (a->n == a->cap) {
            const int cap = a->cap ? 2 * a->cap : 8;
            att_t *na = (att_t *)realloc(a->a, sizeof(att_t) * cap);
            if (!na) return NC3_ENOMEM;
            a->a = na;
            a->cap = cap;
        }
        k = a->n++;
        memset(&a->a[k], 0, sizeof(att_t));
        strcpy(a->a[k].name, name);
    }
    att_t *t = &a->a[k];
    free(t->data);
    t->type = type;
    t->n = n;
    t->data = (unsigned char *)malloc(n * tsize(type) + 1);
    if (!t->data) return NC3_ENOMEM;
    if (n) memcpy(t->data, values, n * tsize(type));
    return NC3_NOERR;
}

/* ---- header reading --------------------------------------------------- */

typedef struct {
    const unsigned char *p, *end;
    int bad;
} rd_t;

static uint32_t rd_u32(rd_t *r) {
    if (r->p + 4 > r->end) {
        r->bad = 1;
        return 0;
    }
    const uint32_t v = ((uint32_t)r->p[0] << 24) | ((uint32_t)r->p[1] << 16) | ((uint32_t)r->p[2] << 8) |
                       (uint32_t)r->p[3];
    r->p += 4;
    return v;
}

static uint64_t rd_u64(rd_t *r) {
    const uint64_t hi = rd_u32(r);
    return (hi << 32) | rd_u32(r);
}

static void rd_name(rd_t *r, char *name) {
    const uint32_t n = rd_u32(r);
    if (r->bad || n >= MAXNAME || r->p + pad4(n) > r->end) {
        r->bad = 1;
        name[0] = 0;
        return;
    }
    memcpy(name, r->p, n);
    name[n] = 0;
    r->p += pad4(n);
}

static int rd_atts(rd_t *r, atts_t *a) {
    const uint32_t tag = rd_u32(r), n = rd_u32(r);
    if (r->bad || (tag != 0 && tag != 0x0C) || (tag == 0 && n != 0)) return NC3_ENOTNC;
    for (uint32_t i = 0; i < n && !r->bad; i++) {
        char name[MAXNAME];
        rd_name(r, name);
        const int type = (int)rd_u32(r);
        const uint32_t cnt = rd_u32(r);
        const size_t s = tsize(type);
        if (r->bad || !s || r->p + pad4(cnt * s) > r->end) return NC3_ENOTNC;
        unsigned char *buf = (unsigned char *)malloc(cnt * s + 1);
        if (!buf) return NC3_ENOMEM;
        memcpy(buf, r->p, cnt * s);
        swap_items(buf, cnt, s);
        r->p += pad4(cnt * s);
        const int rc = add_att(a, name, type, cnt, buf);
        free(buf);
        if (rc) return rc;
    }
    return r->bad ? NC3_ENOTNC : NC3_NOERR;
}

static void layout_records(file_t *f) {
    int nrec = 0, last = -1;
    f->recsize = 0;
    for (int v = 0; v < f->nvars; v++)
        if (f->vars[v].is_rec) {
            nrec++;
            last = v;
            f->recsize += f->vars[v].vsize;
        }
    if (nrec == 1) {  // a lone record variable is not padded
        const var_t *x = &f->vars[last];
        uint64_t n = tsize(x->type);
        for (int d = 1; d < x->ndims; d++) n *= f->dims[x->dimids[d]].len;
        f->recsize = n;
    }
}

static int parse_header(file_t *f, const unsigned char *buf, size_t hn);

/* The header is read in a buffer that doubles until it holds the whole
 * header (a few KB for the files Cliffs writes; the data is not read). */
static int read_header(file_t *f) {
    if (fseek(f->fp, 0, SEEK_END)) return NC3_EIO;
    const long fsz = ftell(f->fp);
    if (fsz < 32) return NC3_ENOTNC;
    for (size_t hn = 64 * 1024;; hn *= 2) {
        if (hn > (size_t)fsz) hn = (size_t)fsz;
        unsigned char *buf = (unsigned char *)malloc(hn);
        if (!buf) return NC3_ENOMEM;
        rewind(f->fp);
        if (fread(buf, 1, hn, f->fp) != hn) {
            free(buf);
            return NC3_EIO;
        }
        const int rc = parse_header(f, buf, hn);
        free(buf);
        if (rc == BADMAGIC) return NC3_ENOTNC;
        if (rc != NC3_ENOTNC || hn == (size_t)fsz) return rc;
        // possibly a header longer than the buffer: start over with a larger one
        for (int v = 0; v < f->nvars; v++) free_atts(&f->vars[v].atts);
        free(f->vars);
        free(f->dims);
        free_atts(&f->gatts);
        f->vars = NULL;
        f->dims = NULL;
        f->nvars = f->ndims = 0;
        f->unlim = -1;
    }
}

static int parse_header(file_t *f, const unsigned char *buf, size_t hn) {
    rd_t r = {buf, buf + hn, 0};
    int rc = NC3_NOERR;
    if (memcmp(buf, "CDF", 3) != 0 || (buf[3] != 1 && buf[3] != 2)) return BADMAGIC;
    f->version = buf[3];
    r.p += 4;
    const uint32_t nrecs = rd_u32(&r);
    f->numrecs = nrecs == 0xFFFFFFFFu ? 0 : nrecs;
    uint32_t tag = rd_u32(&r), n = rd_u32(&r);
    if (tag != 0 && tag != 0x0A) rc = NC3_ENOTNC;
    if (!rc && n > MAXDIMS) rc = NC3_EMAXDIMS;
    if (!rc && n) {
        f->dims = (dim_t *)calloc(n, sizeof(dim_t));
        if (!f->dims) rc = NC3_ENOMEM;
    }
    for (uint32_t i = 0; !rc && i < n; i++) {
        rd_name(&r, f->dims[i].name);
        f->dims[i].len = rd_u32(&r);
        if (f->dims[i].len == 0) f->unlim = (int)i;
    }
    f->ndims = (int)n;
    if (!rc) rc = rd_atts(&r, &f->gatts);
    if (!rc) {
        tag = rd_u32(&r);
        n = rd_u32(&r);
        if (tag != 0 && tag != 0x0B) rc = NC3_ENOTNC;
    }
    if (!rc && n) {
        f->vars = (var_t *)calloc(n, sizeof(var_t));
        if (!f->vars) rc = NC3_ENOMEM;
    }
    for (uint32_t i = 0; !rc && i < n; i++) {
        var_t *v = &f->vars[i];
        f->nvars = (int)i + 1;
        rd_name(&r, v->name);
        v->ndims = (int)rd_u32(&r);
        if (v->ndims > MAXVARDIMS) {
            rc = NC3_EMAXDIMS;
            break;
        }
        for (int d = 0; d < v->ndims; d++) {
            v->dimids[d] = (int)rd_u32(&r);
            if (v->dimids[d] < 0 || v->dimids[d] >= f->ndims) rc = NC3_ENOTNC;
        }
        if (!rc) rc = rd_atts(&r, &v->atts);
        v->type = (int)rd_u32(&r);
        v->vsize = rd_u32(&r);
        v->begin = f->version == 2 ? rd_u64(&r) : rd_u32(&r);
        if (!tsize(v->type)) rc = NC3_EBADTYPE;
        v->is_rec = v->ndims > 0 && v->dimids[0] == f->unlim;
        // recompute the slab size: vsize saturates for variables > 4 GB
        uint64_t s = tsize(v->type);
        for (int d = v->is_rec ? 1 : 0; d < v->ndims && !rc; d++) s *= f->dims[v->dimids[d]].len;
        v->vsize = pad4(s);
    }
    if (!rc && r.bad) rc = NC3_ENOTNC;
    if (rc) return rc;
    layout_records(f);
    f->recbegin = 0;
    for (int v = 0; v < f->nvars; v++)
        if (f->vars[v].is_rec && (f->recbegin == 0 || f->vars[v].begin < f->recbegin))
            f->recbegin = f->vars[v].begin;
    return NC3_NOERR;
}

/* ---- header writing --------------------------------------------------- */

typedef struct {
    unsigned char *p;
    size_t n, cap;
    int oom;
} wr_t;

static void wr_bytes(wr_t *w, const void *src, size_t n) {
    if (w->n + n > w->cap) {
        size_t cap = w->cap ? w->cap : 1024;
        while (cap < w->n + n) cap *= 2;
        unsigned char *np = (unsigned char *)realloc(w->p, cap);
        if (!np) {
            w->oom = 1;
            return;
        }
        w->p = np;
        w->cap = cap;
    }
    memcpy(w->p + w->n, src, n);
    w->n += n;
}

static void wr_u32(wr_t *w, uint32_t v) {
    const unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
                                (unsigned char)v};
    wr_bytes(w, b, 4);
}

static void wr_u64(wr_t *w, uint64_t v) {
    wr_u32(w, (uint32_t)(v >> 32));
    wr_u32(w, (uint32_t)v);
}

static void wr_pad(wr_t *w) {
    static const unsigned char z[4] = {0, 0, 0, 0};
    wr_bytes(w, z, pad4(w->n) - w->n);
}

static void wr_name(wr_t *w, const char *s) {
    const size_t n = strlen(s);
    wr_u32(w, (uint32_t)n);
    wr_bytes(w, s, n);
    wr_pad(w);
}

static void wr_atts(wr_t *w, const atts_t *a) {
    wr_u32(w, a->n ? 0x0C : 0);
    wr_u32(w, (uint32_t)a->n);
    for (int i = 0; i < a->n; i++) {
        const att_t *t = &a->a[i];
        const size_t s = tsize(t->type);
        wr_name(w, t->name);
        wr_u32(w, (uint32_t)t->type);
        wr_u32(w, (uint32_t)t->n);
        unsigned char *tmp = (unsigned char *)malloc(t->n * s + 1);
        if (!tmp) {
            w->oom = 1;
            return;
        }
        memcpy(tmp, t->data, t->n * s);
        swap_items(tmp, t->n, s);
        wr_bytes(w, tmp, t->n * s);
        free(tmp);
        wr_pad(w);
    }
}

static void build_header(const file_t *f, wr_t *w) {
    const unsigned char magic[4] = {'C', 'D', 'F', (unsigned char)f->version};
    wr_bytes(w, magic, 4);
    wr_u32(w, (uint32_t)f->numrecs);
    wr_u32(w, f->ndims ? 0x0A : 0);
    wr_u32(w, (uint32_t)f->ndims);
    for (int i = 0; i < f->ndims; i++) {
        wr_name(w, f->dims[i].name);
        wr_u32(w, (uint32_t)f->dims[i].len);
    }
    wr_atts(w, &f->gatts);
    wr_u32(w, f->nvars ? 0x0B : 0);
    wr_u32(w, (uint32_t)f->nvars);
    for (int v = 0; v < f->nvars; v++) {
        const var_t *x = &f->vars[v];
        wr_name(w, x->name);
        wr_u32(w, (uint32_t)x->ndims);
        for (int d = 0; d < x->ndims; d++) wr_u32(w, (uint32_t)x->dimids[d]);
        wr_atts(w, &x->atts);
        wr_u32(w, (uint32_t)x->type);
        wr_u32(w, x->vsize > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)x->vsize);
        if (f->version == 2)
            wr_u64(w, x->begin);
        else
            wr_u32(w, (uint32_t)x->begin);
    }
}

/* ---- fill values ------------------------------------------------------ */

static void fill_item(int type, unsigned char *out) {
    switch (type) {
        case NC3_BYTE: *(int8_t *)out = -127; break;
        case NC3_CHAR: *out = 0; break;
        case NC3_SHORT: *(int16_t *)out = -32767; break;
        case NC3_INT: *(int32_t *)out = -2147483647; break;
        case NC3_FLOAT: *(float *)out = 9.9692099683868690e+36f; break;
        case NC3_DOUBLE: *(double *)out = 9.9692099683868690e+36; break;
    }
}

/* write `bytes` bytes of var type `type` fill at file offset off */
static int write_fill(file_t *f, uint64_t off, uint64_t bytes, int type) {
    const size_t s = tsize(type);
    unsigned char chunk[4096];
    for (size_t k = 0; k + s <= sizeof(chunk); k += s) fill_item(type, chunk + k);
    swap_items(chunk, sizeof(chunk) / s, s);
    if (fseek(f->fp, (long)off, SEEK_SET)) return NC3_EIO;
    while (bytes) {
        const size_t n = bytes < sizeof(chunk) ? (size_t)bytes : sizeof(chunk);
        if (fwrite(chunk, 1, n, f->fp) != n) return NC3_EIO;
        bytes -= n;
    }
    return NC3_NOERR;
}

static int grow_records(file_t *f, size_t nrecs) {
    for (size_t r = f->numrecs; r < nrecs; r++)
        for (int v = 0; v < f->nvars; v++) {
            const var_t *x = &f->vars[v];
            if (!x->is_rec) continue;
            // a lone record variable's slab is unpadded: recsize < vsize
            const uint64_t n = x->vsize < f->recsize ? x->vsize : f->recsize;
            const int rc = write_fill(f, x->begin + r * f->recsize, n, x->type);
            if (rc) return rc;
        }
    if (nrecs > f->numrecs) f->numrecs = nrecs;
    return NC3_NOERR;
}

/* ---- the C API -------------------------------------------------------- */

int nc3_create(const char *path, int cmode, int *ncid) {
    if (!path || !ncid) return NC3_EINVAL;
    if (cmode & NC3_NOCLOBBER) {
        FILE *t = fopen(path, "rb");
        if (t) {
            fclose(t);
            return NC3_EEXIST;
        }
    }
    int rc = new_file(ncid);
    if (rc) return rc;
    file_t *f = &files[*ncid];
    f->fp = fopen(path, "w+b");
    if (!f->fp) {
        free_file(f);
        return NC3_EPERM;
    }
    f->writable = 1;
    f->define = 1;
    f->version = (cmode & NC3_64BIT_OFFSET) ? 2 : 1;
    return NC3_NOERR;
}

int nc3_open(const char *path, int mode, int *ncid) {
    if (!path || !ncid) return NC3_EINVAL;
    int rc = new_file(ncid);
    if (rc) return rc;
    file_t *f = &files[*ncid];
    f->writable = (mode & NC3_WRITE) != 0;
    f->fp = fopen(path, f->writable ? "r+b" : "rb");
    if (!f->fp) {
        free_file(f);
        return NC3_EPERM;  // the library reports a system error here; any nonzero code will do
    }
    rc = read_header(f);
    if (rc) free_file(f);
    return rc;
}

int nc3_def_dim(int ncid, const char *name, size_t len, int *dimid) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (!f->define) return NC3_ENOTINDEFINE;
    if (!name || strlen(name) >= MAXNAME) return NC3_EMAXNAME;
    if (len == 0 && f->unlim >= 0) return NC3_EUNLIMIT;
    for (int i = 0; i < f->ndims; i++)
        if (strcmp(f->dims[i].name, name) == 0) return NC3_ENAMEINUSE;
    if (f->ndims >= MAXDIMS) return NC3_EMAXDIMS;
    dim_t *nd = (dim_t *)realloc(f->dims, sizeof(dim_t) * (f->ndims + 1));
    if (!nd) return NC3_ENOMEM;
    f->dims = nd;
    memset(&nd[f->ndims], 0, sizeof(dim_t));
    strcpy(nd[f->ndims].name, name);
    nd[f->ndims].len = len;
    if (len == 0) f->unlim = f->ndims;
    if (dimid) *dimid = f->ndims;
    f->ndims++;
    return NC3_NOERR;
}

int nc3_def_var(int ncid, const char *name, int type, int ndims, const int *dimids, int *varid) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (!f->define) return NC3_ENOTINDEFINE;
    if (!name || strlen(name) >= MAXNAME) return NC3_EMAXNAME;
    if (!tsize(type)) return NC3_EBADTYPE;
    if (ndims < 0 || ndims > MAXVARDIMS || (ndims > 0 && !dimids)) return NC3_EINVAL;
    for (int v = 0; v < f->nvars; v++)
        if (strcmp(f->vars[v].name, name) == 0) return NC3_ENAMEINUSE;
    for (int d = 0; d < ndims; d++) {
        if (dimids[d] < 0 || dimids[d] >= f->ndims) return NC3_EBADDIM;
        if (d > 0 && dimids[d] == f->unlim) return NC3_EUNLIMPOS;
    }
    var_t *nv = (var_t *)realloc(f->vars, sizeof(var_t) * (f->nvars + 1));
    if (!nv) return NC3_ENOMEM;
    f->vars = nv;
    var_t *v = &nv[f->nvars];
    memset(v, 0, sizeof(*v));
    strcpy(v->name, name);
    v->type = type;
    v->ndims = ndims;
    for (int d = 0; d < ndims; d++) v->dimids[d] = dimids[d];
    v->is_rec = ndims > 0 && dimids[0] == f->unlim;
    if (varid) *varid = f->nvars;
    f->nvars++;
    return NC3_NOERR;
}

int nc3_put_att(int ncid, int varid, const char *name, int type, size_t n, const void *values) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (!f->define) return NC3_ENOTINDEFINE;
    atts_t *a = atts_of(f, varid);
    if (!a) return NC3_ENOTVAR;
    if (!name) return NC3_EINVAL;
    if (n && !values) return NC3_EINVAL;
    return add_att(a, name, type, n, values);
}

int nc3_copy_att(int ncid_in, int varid_in, const char *name, int ncid_out, int varid_out) {
    file_t *fi = get(ncid_in);
    if (!fi || !get(ncid_out)) return NC3_EBADID;
    atts_t *a = atts_of(fi, varid_in);
    if (!a) return NC3_ENOTVAR;
    const int k = find_att(a, name);
    if (k < 0) return NC3_ENOTATT;
    // copy first: add_att may move the source array when both are the same list
    const att_t src = a->a[k];
    unsigned char *tmp = (unsigned char *)malloc(src.n * tsize(src.type) + 1);
    if (!tmp) return NC3_ENOMEM;
    memcpy(tmp, src.data, src.n * tsize(src.type));
    const int rc = nc3_put_att(ncid_out, varid_out, name, src.type, src.n, tmp);
    free(tmp);
    return rc;
}

int nc3_enddef(int ncid) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (!f->define) return NC3_ENOTINDEFINE;
    for (int v = 0; v < f->nvars; v++) {
        var_t *x = &f->vars[v];
        uint64_t s = tsize(x->type);
        for (int d = x->is_rec ? 1 : 0; d < x->ndims; d++) s *= f->dims[x->dimids[d]].len;
        x->vsize = pad4(s);
    }
    wr_t w = {NULL, 0, 0, 0};
    build_header(f, &w);  // offsets are placeholders; the header's size does not depend on them
    if (w.oom) {
        free(w.p);
        return NC3_ENOMEM;
    }
    uint64_t off = w.n;
    for (int v = 0; v < f->nvars; v++)
        if (!f->vars[v].is_rec) {
            f->vars[v].begin = off;
            off += f->vars[v].vsize;
        }
    f->recbegin = off;
    for (int v = 0; v < f->nvars; v++)
        if (f->vars[v].is_rec) {
            f->vars[v].begin = off;
            off += f->vars[v].vsize;
        }
    if (f->version == 1 && off > 0x7FFFFFFFull) {
        free(w.p);
        return NC3_EINVAL;  // needs the 64-bit-offset format
    }
    layout_records(f);
    w.n = 0;
    build_header(f, &w);
    int rc = w.oom ? NC3_ENOMEM : NC3_NOERR;
    if (!rc && (fseek(f->fp, 0, SEEK_SET) || fwrite(w.p, 1, w.n, f->fp) != w.n)) rc = NC3_EIO;
    free(w.p);
    for (int v = 0; v < f->nvars && !rc; v++)
        if (!f->vars[v].is_rec) rc = write_fill(f, f->vars[v].begin, f->vars[v].vsize, f->vars[v].type);
    if (!rc) f->define = 0;
    return rc;
}

int nc3_sync(int ncid) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (f->define) return NC3_EINDEFINE;
    if (f->writable) {
        wr_t w = {NULL, 0, 0, 0};
        wr_u32(&w, (uint32_t)f->numrecs);
        int rc = w.oom ? NC3_ENOMEM : NC3_NOERR;
        if (!rc && (fseek(f->fp, 4, SEEK_SET) || fwrite(w.p, 1, 4, f->fp) != 4)) rc = NC3_EIO;
        free(w.p);
        if (rc) return rc;
        if (fflush(f->fp)) return NC3_EIO;
    }
    return NC3_NOERR;
}

int nc3_close(int ncid) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    int rc = NC3_NOERR;
    if (f->define) rc = nc3_enddef(ncid);
    if (!rc && f->writable) rc = nc3_sync(ncid);
    free_file(f);
    return rc;
}

int nc3_inq(int ncid, int *ndims, int *nvars, int *ngatts, int *unlimdimid) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (ndims) *ndims = f->ndims;
    if (nvars) *nvars = f->nvars;
    if (ngatts) *ngatts = f->gatts.n;
    if (unlimdimid) *unlimdimid = f->unlim;
    return NC3_NOERR;
}

int nc3_inq_dim(int ncid, int dimid, char *name, size_t *len) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (dimid < 0 || dimid >= f->ndims) return NC3_EBADDIM;
    if (name) strcpy(name, f->dims[dimid].name);
    if (len) *len = dimid == f->unlim ? f->numrecs : f->dims[dimid].len;
    return NC3_NOERR;
}

int nc3_inq_var(int ncid, int varid, char *name, int *type, int *ndims, int *dimids, int *natts) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (varid < 0 || varid >= f->nvars) return NC3_ENOTVAR;
    const var_t *v = &f->vars[varid];
    if (name) strcpy(name, v->name);
    if (type) *type = v->type;
    if (ndims) *ndims = v->ndims;
    if (dimids)
        for (int d = 0; d < v->ndims; d++) dimids[d] = v->dimids[d];
    if (natts) *natts = v->atts.n;
    return NC3_NOERR;
}

int nc3_inq_attname(int ncid, int varid, int attnum, char *name) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    atts_t *a = atts_of(f, varid);
    if (!a) return NC3_ENOTVAR;
    if (attnum < 0 || attnum >= a->n) return NC3_ENOTATT;
    if (name) strcpy(name, a->a[attnum].name);
    return NC3_NOERR;
}

int nc3_inq_att(int ncid, int varid, const char *name, int *type, size_t *n) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    atts_t *a = atts_of(f, varid);
    if (!a) return NC3_ENOTVAR;
    const int k = find_att(a, name);
    if (k < 0) return NC3_ENOTATT;
    if (type) *type = a->a[k].type;
    if (n) *n = a->a[k].n;
    return NC3_NOERR;
}

int nc3_get_att(int ncid, int varid, const char *name, void *values) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    atts_t *a = atts_of(f, varid);
    if (!a) return NC3_ENOTVAR;
    const int k = find_att(a, name);
    if (k < 0) return NC3_ENOTATT;
    memcpy(values, a->a[k].data, a->a[k].n * tsize(a->a[k].type));
    return NC3_NOERR;
}

/* file item (host order after swap) -> double */
static double item_to_double(int type, const unsigned char *p) {
    switch (type) {
        case NC3_BYTE: return (double)*(const int8_t *)p;
        case NC3_CHAR: return (double)*p;
        case NC3_SHORT: return (double)*(const int16_t *)p;
        case NC3_INT: return (double)*(const int32_t *)p;
        case NC3_FLOAT: return (double)*(const float *)p;
        case NC3_DOUBLE: return *(const double *)p;
    }
    return 0.0;
}

static void double_to_item(int type, double x, unsigned char *p) {
    switch (type) {
        case NC3_BYTE: *(int8_t *)p = (int8_t)x; break;
        case NC3_CHAR: *p = (unsigned char)x; break;
        case NC3_SHORT: *(int16_t *)p = (int16_t)x; break;
        case NC3_INT: *(int32_t *)p = (int32_t)x; break;
        case NC3_FLOAT: *(float *)p = (float)x; break;
        case NC3_DOUBLE: *(double *)p = x; break;
    }
}

/* Visit the contiguous runs of a hyperslab: fn(file offset, element offset
 * in the caller's array, run length). */
typedef int (*run_fn)(file_t *f, const var_t *v, uint64_t off, size_t at, size_t n, void *ctx);

static int for_runs(file_t *f, const var_t *v, const size_t *start, const size_t *count, run_fn fn,
                    void *ctx) {
    const int nd = v->ndims;
    const size_t s = tsize(v->type);
    if (nd == 0) return fn(f, v, v->begin, 0, 1, ctx);
    if (v->is_rec && nd == 1) {  // one value per record: records are recsize apart
        for (size_t r = 0; r < count[0]; r++) {
            const int rc = fn(f, v, v->begin + (start[0] + r) * f->recsize, r, 1, ctx);
            if (rc) return rc;
        }
        return NC3_NOERR;
    }
    size_t len[MAXVARDIMS];
    for (int d = 0; d < nd; d++) len[d] = f->dims[v->dimids[d]].len;
    // element strides within one variable (or one record slab)
    uint64_t stride[MAXVARDIMS];
    stride[nd - 1] = 1;
    for (int d = nd - 2; d >= 0; d--) stride[d] = stride[d + 1] * (v->is_rec && d == 0 ? 1 : len[d + 1]);
    for (int d = 0; d < nd; d++)
        if (count[d] == 0) return NC3_NOERR;
    size_t idx[MAXVARDIMS] = {0};
    size_t at = 0;
    const size_t run = count[nd - 1];
    for (;;) {
        uint64_t off = v->begin;
        for (int d = 0; d < nd; d++) {
            const uint64_t i = start[d] + idx[d];
            if (v->is_rec && d == 0)
                off += i * f->recsize;
            else
                off += i * stride[d] * s;
        }
        const int rc = fn(f, v, off, at, run, ctx);
        if (rc) return rc;
        at += run;
        int d = nd - 2;
        for (; d >= 0; d--) {
            if (++idx[d] < count[d]) break;
            idx[d] = 0;
        }
        if (d < 0) break;
    }
    return NC3_NOERR;
}

typedef struct {
    void *mem;
    int memtype;
    unsigned char *buf;
    size_t cap;
} io_ctx;

static int ensure(io_ctx *c, size_t bytes) {
    if (bytes <= c->cap) return NC3_NOERR;
    unsigned char *nb = (unsigned char *)realloc(c->buf, bytes);
    if (!nb) return NC3_ENOMEM;
    c->buf = nb;
    c->cap = bytes;
    return NC3_NOERR;
}

static int read_run(file_t *f, const var_t *v, uint64_t off, size_t at, size_t n, void *vctx) {
    io_ctx *c = (io_ctx *)vctx;
    const size_t s = tsize(v->type);
    int rc = ensure(c, n * s);
    if (rc) return rc;
    if (fseek(f->fp, (long)off, SEEK_SET) || fread(c->buf, s, n, f->fp) != n) return NC3_EIO;
    swap_items(c->buf, n, s);
    if (c->memtype == NC3_DOUBLE) {
        double *o = (double *)c->mem + at;
        for (size_t k = 0; k < n; k++) o[k] = item_to_double(v->type, c->buf + k * s);
    } else {
        float *o = (float *)c->mem + at;
        for (size_t k = 0; k < n; k++) o[k] = (float)item_to_double(v->type, c->buf + k * s);
    }
    return NC3_NOERR;
}

static int write_run(file_t *f, const var_t *v, uint64_t off, size_t at, size_t n, void *vctx) {
    io_ctx *c = (io_ctx *)vctx;
    const size_t s = tsize(v->type);
    int rc = ensure(c, n * s);
    if (rc) return rc;
    if (c->memtype == NC3_DOUBLE) {
        const double *in = (const double *)c->mem + at;
        for (size_t k = 0; k < n; k++) double_to_item(v->type, in[k], c->buf + k * s);
    } else {
        const float *in = (const float *)c->mem + at;
        for (size_t k = 0; k < n; k++) double_to_item(v->type, (double)in[k], c->buf + k * s);
    }
    swap_items(c->buf, n, s);
    if (fseek(f->fp, (long)off, SEEK_SET) || fwrite(c->buf, s, n, f->fp) != n) return NC3_EIO;
    return NC3_NOERR;
}

static int slab(file_t *f, const var_t *v, const size_t *start, const size_t *count, size_t *st,
                size_t *ct, int writing) {
    for (int d = 0; d < v->ndims; d++) {
        const size_t len = (v->is_rec && d == 0) ? f->numrecs : f->dims[v->dimids[d]].len;
        st[d] = start ? start[d] : 0;
        ct[d] = count ? count[d] : len - st[d];
        if (v->is_rec && d == 0 && writing) continue;  // records may be appended
        if (st[d] > len || (st[d] == len && ct[d] > 0)) return NC3_EINVALCOORDS;
        if (st[d] + ct[d] > len) return NC3_EEDGE;
    }
    return NC3_NOERR;
}

int nc3_get_vara(int ncid, int varid, const size_t *start, const size_t *count, void *out, int memtype) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (f->define) return NC3_EINDEFINE;
    if (varid < 0 || varid >= f->nvars) return NC3_ENOTVAR;
    if (memtype != NC3_FLOAT && memtype != NC3_DOUBLE) return NC3_EBADTYPE;
    const var_t *v = &f->vars[varid];
    if (v->type == NC3_CHAR) return NC3_ECHAR;
    size_t st[MAXVARDIMS], ct[MAXVARDIMS];
    int rc = slab(f, v, start, count, st, ct, 0);
    if (rc) return rc;
    io_ctx c = {out, memtype, NULL, 0};
    rc = for_runs(f, v, st, ct, read_run, &c);
    free(c.buf);
    return rc;
}

int nc3_put_vara(int ncid, int varid, const size_t *start, const size_t *count, const void *in,
                 int memtype) {
    file_t *f = get(ncid);
    if (!f) return NC3_EBADID;
    if (f->define) return NC3_EINDEFINE;
    if (!f->writable) return NC3_EPERM;
    if (varid < 0 || varid >= f->nvars) return NC3_ENOTVAR;
    if (memtype != NC3_FLOAT && memtype != NC3_DOUBLE) return NC3_EBADTYPE;
    const var_t *v = &f->vars[varid];
    if (v->type == NC3_CHAR) return NC3_ECHAR;
    size_t st[MAXVARDIMS], ct[MAXVARDIMS];
    int rc = slab(f, v, start, count, st, ct, 1);
    if (rc) return rc;
    if (v->is_rec && st[0] + ct[0] > f->numrecs) {
        rc = grow_records(f, st[0] + ct[0]);
        if (rc) return rc;
    }
    io_ctx c = {(void *)in, memtype, NULL, 0};
    rc = for_runs(f, v, st, ct, write_run, &c);
    free(c.buf);
    return rc;
}

const char *nc3_strerror(int status) {
    switch (status) {
        case NC3_NOERR: return "No error";
        case NC3_EBADID: return "NetCDF: Not a valid ID";
        case NC3_ENFILE: return "NetCDF: Too many files open";
        case NC3_EEXIST: return "NetCDF: File exists && NC_NOCLOBBER";
        case NC3_EINVAL: return "NetCDF: Invalid argument";
        case NC3_EPERM: return "NetCDF: Write to read only, or file cannot be opened";
        case NC3_ENOTINDEFINE: return "NetCDF: Operation not allowed in data mode";
        case NC3_EINDEFINE: return "NetCDF: Operation not allowed in define mode";
        case NC3_EINVALCOORDS: return "NetCDF: Index exceeds dimension bound";
        case NC3_EMAXDIMS: return "NetCDF: NC_MAX_DIMS exceeded";
        case NC3_ENAMEINUSE: return "NetCDF: String match to name in use";
        case NC3_ENOTATT: return "NetCDF: Attribute not found";
        case NC3_EMAXATTS: return "NetCDF: NC_MAX_ATTRS exceeded";
        case NC3_EBADTYPE: return "NetCDF: Not a valid data type or _FillValue type mismatch";
        case NC3_EBADDIM: return "NetCDF: Invalid dimension ID or name";
        case NC3_EUNLIMPOS: return "NetCDF: NC_UNLIMITED in the wrong index";
        case NC3_EMAXVARS: return "NetCDF: NC_MAX_VARS exceeded";
        case NC3_ENOTVAR: return "NetCDF: Variable not found";
        case NC3_EGLOBAL: return "NetCDF: Action prohibited on NC_GLOBAL varid";
        case NC3_ENOTNC: return "NetCDF: Unknown file format";
        case NC3_EMAXNAME: return "NetCDF: NC_MAX_NAME exceeded";
        case NC3_EUNLIMIT: return "NetCDF: NC_UNLIMITED size already in use";
        case NC3_ECHAR: return "NetCDF: Attempt to convert between text & numbers";
        case NC3_EEDGE: return "NetCDF: Start+count exceeds dimension bound";
        case NC3_ERANGE: return "NetCDF: Numeric conversion not representable";
        case NC3_ENOMEM: return "NetCDF: Memory allocation (malloc) failure";
        case NC3_EIO: return "NetCDF: I/O failure";
    }
    return "NetCDF: Unknown error";
}

/* ---- Fortran-77 entry points -------------------------------------------
 * Fortran conventions: every argument by reference, CHARACTER lengths
 * appended as hidden size_t arguments, ids and indices 1-based, NF_GLOBAL =
 * 0, dimensions listed fastest first (the reverse of C).  A CHARACTER
 * function returns through a hidden (buffer, length) pair in front. */

static void f2c_str(const char *s, size_t len, char *out, size_t cap) {
    while (len > 0 && (s[len - 1] == ' ' || s[len - 1] == 0)) len--;
    if (len >= cap) len = cap - 1;
    memcpy(out, s, len);
    out[len] = 0;
}

static void c2f_str(const char *s, char *out, size_t len) {
    size_t n = strlen(s);
    if (n > len) n = len;
    memcpy(out, s, n);
    memset(out + n, ' ', len - n);
}

static int f_ndims(int ncid, int varid) {
    int nd = 0;
    if (nc3_inq_var(ncid, varid, NULL, NULL, &nd, NULL, NULL)) return -1;
    return nd;
}

int nf_create_(const char *path, const int *cmode, int *ncid, size_t len) {
    char p[4096];
    f2c_str(path, len, p, sizeof(p));
    return nc3_create(p, *cmode, ncid);
}

int nf_open_(const char *path, const int *mode, int *ncid, size_t len) {
    char p[4096];
    f2c_str(path, len, p, sizeof(p));
    return nc3_open(p, *mode, ncid);
}

int nf_close_(const int *ncid) { return nc3_close(*ncid); }
int nf_sync_(const int *ncid) { return nc3_sync(*ncid); }
int nf_enddef_(const int *ncid) { return nc3_enddef(*ncid); }

int nf_inq_(const int *ncid, int *ndims, int *nvars, int *ngatts, int *unlim) {
    int u = -1;
    const int rc = nc3_inq(*ncid, ndims, nvars, ngatts, &u);
    if (!rc) *unlim = u < 0 ? -1 : u + 1;
    return rc;
}

int nf_inq_dim_(const int *ncid, const int *dimid, char *name, int *len, size_t name_len) {
    char n[MAXNAME];
    size_t l = 0;
    const int rc = nc3_inq_dim(*ncid, *dimid - 1, n, &l);
    if (!rc) {
        c2f_str(n, name, name_len);
        *len = (int)l;
    }
    return rc;
}

int nf_inq_dimlen_(const int *ncid, const int *dimid, int *len) {
    size_t l = 0;
    const int rc = nc3_inq_dim(*ncid, *dimid - 1, NULL, &l);
    if (!rc) *len = (int)l;
    return rc;
}

int nf_inq_varname_(const int *ncid, const int *varid, char *name, size_t name_len) {
    char n[MAXNAME];
    const int rc = nc3_inq_var(*ncid, *varid - 1, n, NULL, NULL, NULL, NULL);
    if (!rc) c2f_str(n, name, name_len);
    return rc;
}

int nf_inq_vartype_(const int *ncid, const int *varid, int *type) {
    return nc3_inq_var(*ncid, *varid - 1, NULL, type, NULL, NULL, NULL);
}

int nf_inq_var_(const int *ncid, const int *varid, char *name, int *type, int *ndims, int *dimids,
                int *natts, size_t name_len) {
    char n[MAXNAME];
    int d[MAXVARDIMS], nd = 0;
    const int rc = nc3_inq_var(*ncid, *varid - 1, n, type, &nd, d, natts);
    if (!rc) {
        c2f_str(n, name, name_len);
        *ndims = nd;
        for (int k = 0; k < nd; k++) dimids[k] = d[nd - 1 - k] + 1;
    }
    return rc;
}

int nf_inq_attname_(const int *ncid, const int *varid, const int *attnum, char *name, size_t name_len) {
    char n[MAXNAME];
    const int rc = nc3_inq_attname(*ncid, *varid - 1, *attnum - 1, n);
    if (!rc) c2f_str(n, name, name_len);
    return rc;
}

int nf_def_dim_(const int *ncid, const char *name, const int *len, int *dimid, size_t name_len) {
    char n[MAXNAME];
    f2c_str(name, name_len, n, sizeof(n));
    if (*len < 0) return NC3_EINVAL;
    int d = -1;
    const int rc = nc3_def_dim(*ncid, n, (size_t)*len, &d);
    if (!rc) *dimid = d + 1;
    return rc;
}

int nf_def_var_(const int *ncid, const char *name, const int *type, const int *ndims, const int *dimids,
                int *varid, size_t name_len) {
    char n[MAXNAME];
    f2c_str(name, name_len, n, sizeof(n));
    const int nd = *ndims;
    if (nd < 0 || nd > MAXVARDIMS) return NC3_EINVAL;
    int d[MAXVARDIMS];
    for (int k = 0; k < nd; k++) d[k] = dimids[nd - 1 - k] - 1;
    int v = -1;
    const int rc = nc3_def_var(*ncid, n, *type, nd, d, &v);
    if (!rc) *varid = v + 1;
    return rc;
}

int nf_put_att_text_(const int *ncid, const int *varid, const char *name, const int *len, const char *text,
                     size_t name_len, size_t text_len) {
    char n[MAXNAME];
    f2c_str(name, name_len, n, sizeof(n));
    if (*len < 0 || (size_t)*len > text_len) return NC3_EINVAL;
    return nc3_put_att(*ncid, *varid - 1, n, NC3_CHAR, (size_t)*len, text);
}

int nf_copy_att_(const int *ncid_in, const int *varid_in, const char *name, const int *ncid_out,
                 const int *varid_out, size_t name_len) {
    char n[MAXNAME];
    f2c_str(name, name_len, n, sizeof(n));
    return nc3_copy_att(*ncid_in, *varid_in - 1, n, *ncid_out, *varid_out - 1);
}

static int f_vara(const int *ncid, const int *varid, const int *start, const int *count, void *data,
                  int memtype, int writing) {
    const int nd = f_ndims(*ncid, *varid - 1);
    if (nd < 0) return get(*ncid) ? NC3_ENOTVAR : NC3_EBADID;
    size_t st[MAXVARDIMS], ct[MAXVARDIMS];
    for (int k = 0; k < nd; k++) {
        if (start[nd - 1 - k] < 1 || count[nd - 1 - k] < 0) return NC3_EINVALCOORDS;
        st[k] = (size_t)(start[nd - 1 - k] - 1);
        ct[k] = (size_t)count[nd - 1 - k];
    }
    return writing ? nc3_put_vara(*ncid, *varid - 1, st, ct, data, memtype)
                   : nc3_get_vara(*ncid, *varid - 1, st, ct, data, memtype);
}

int nf_get_var_double_(const int *ncid, const int *varid, double *out) {
    return nc3_get_vara(*ncid, *varid - 1, NULL, NULL, out, NC3_DOUBLE);
}
int nf_get_var_real_(const int *ncid, const int *varid, float *out) {
    return nc3_get_vara(*ncid, *varid - 1, NULL, NULL, out, NC3_FLOAT);
}
int nf_get_vara_double_(const int *ncid, const int *varid, const int *start, const int *count, double *out) {
    return f_vara(ncid, varid, start, count, out, NC3_DOUBLE, 0);
}
int nf_get_vara_real_(const int *ncid, const int *varid, const int *start, const int *count, float *out) {
    return f_vara(ncid, varid, start, count, out, NC3_FLOAT, 0);
}
int nf_put_var_double_(const int *ncid, const int *varid, const double *in) {
    return nc3_put_vara(*ncid, *varid - 1, NULL, NULL, in, NC3_DOUBLE);
}
int nf_put_var_real_(const int *ncid, const int *varid, const float *in) {
    return nc3_put_vara(*ncid, *varid - 1, NULL, NULL, in, NC3_FLOAT);
}
int nf_put_vara_double_(const int *ncid, const int *varid, const int *start, const int *count,
                        const double *in) {
    return f_vara(ncid, varid, start, count, (void *)in, NC3_DOUBLE, 1);
}
int nf_put_vara_real_(const int *ncid, const int *varid, const int *start, const int *count,
                      const float *in) {
    return f_vara(ncid, varid, start, count, (void *)in, NC3_FLOAT, 1);
}

/* CHARACTER*80 FUNCTION nf_strerror(status) */
void nf_strerror_(char *result, size_t result_len, const int *status) {
    c2f_str(nc3_strerror(*status), result, result_len);
}
