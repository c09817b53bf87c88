/*
 * qrdm_pymod.c -- the Python module "QRDM", drop-in for the reference's
 * CPython binding (QRDM_wrapper.c:1-159): the same five functions with the
 * same argument tuples and return values.
 *
 *   QRDM.QRDM(layout, m, n, A, lda, jpvt, tau, ncols, thres, nb) -> info
 *       QRDM_wrapper.c:71-101; runs dgeqrdm of libqrdm_amd (the GPU path).
 *   QRDM.QRF(layout, m, n, A, lda, tau) -> info
 *       QRDM_wrapper.c:44-68 (LAPACKE_dgeqrf); here the GPU unpivoted QR.
 *   QRDM.DORMQR(layout, m, n, k, A, lda, tau, C, ldc) -> info
 *       QRDM_wrapper.c:104-126 (LAPACKE_dormqr 'L','N'); here on the GPU.
 *   QRDM.QP3(layout, m, n, A, lda, jpvt, tau) -> info
 *       QRDM_wrapper.c:15-41: a forward to the host LAPACKE_dgeqp3 (the
 *       OpenBLAS bundled with scipy, found at first use), as in the reference.
 *   QRDM.init() -> None      QRDM_wrapper.c:9-13.
 *
 * Arrays are used in place through their raw data pointers, interpreted
 * column-major with the given leading dimension, whatever their numpy
 * strides (QRDM_wrapper.c:155-159), so a C-ordered (m,n) array with layout
 * 102 factors its transpose, as the reference does.  The only checks added
 * are ones that cannot change a result: each array must expose a contiguous
 * buffer (the reference would read a non-contiguous one's memory blindly).
 * The GIL is released while the device works.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/qrdm_amd.h"

typedef struct {
  Py_buffer view;
  int held;
} buf_t;

/* out != 0: the library writes the buffer, which must then be writable (a
 * read-only buffer is refused rather than written through). */
static int get_buf_rw(PyObject* o, buf_t* b, const char* name, int out) {
  b->held = 0;
  if (PyObject_GetBuffer(o, &b->view, (out ? PyBUF_WRITABLE : 0) | PyBUF_ANY_CONTIGUOUS) != 0) {
    PyErr_Clear();
    PyErr_Format(PyExc_TypeError, "QRDM: argument '%s' must expose a contiguous%s buffer", name,
                 out ? " writable" : "");
    return -1;
  }
  b->held = 1;
  return 0;
}
#define get_buf(o, b, name) get_buf_rw(o, b, name, 1)
#define get_in_buf(o, b, name) get_buf_rw(o, b, name, 0)

static void rel(buf_t* b) {
  if (b->held) PyBuffer_Release(&b->view);
  b->held = 0;
}

static PyObject* py_init(PyObject* self, PyObject* args) {
  (void)self;
  (void)args;
  Py_RETURN_NONE;
}

static PyObject* py_qrdm(PyObject* self, PyObject* args) {
  (void)self;
  int layout, m, n, lda, nb;
  PyObject *oA, *ojp, *otau, *onc, *oth;
  if (!PyArg_ParseTuple(args, "iiiOiOOOOi", &layout, &m, &n, &oA, &lda, &ojp, &otau, &onc,
                        &oth, &nb))
    return NULL;
  buf_t A, jp, tau, nc, th;
  if (get_buf(oA, &A, "A")) return NULL;
  if (get_buf(ojp, &jp, "jpvt")) { rel(&A); return NULL; }
  if (get_buf(otau, &tau, "tau")) { rel(&A); rel(&jp); return NULL; }
  if (get_buf(onc, &nc, "ncols")) { rel(&A); rel(&jp); rel(&tau); return NULL; }
  if (get_in_buf(oth, &th, "thres")) { rel(&A); rel(&jp); rel(&tau); rel(&nc); return NULL; }
  int out;
  Py_BEGIN_ALLOW_THREADS
  out = dgeqrdm(layout, m, n, (double*)A.view.buf, lda, (int*)jp.view.buf, (double*)tau.view.buf,
                (int*)nc.view.buf, (double*)th.view.buf, nb);
  Py_END_ALLOW_THREADS
  rel(&A); rel(&jp); rel(&tau); rel(&nc); rel(&th);
  return Py_BuildValue("i", out);
}

/* Row-major (101) operands are transposed into column-major scratch and back,
 * as LAPACKE's high-level interface does. */
static double* to_colmajor(const double* a, int rows, int cols, int lda_row) {
  double* t = (double*)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1) * (cols > 0 ? cols : 1));
  if (!t) return NULL;
  for (int i = 0; i < rows; i++)
    for (int j = 0; j < cols; j++) t[i + (size_t)j * rows] = a[(size_t)i * lda_row + j];
  return t;
}

static void from_colmajor(double* a, const double* t, int rows, int cols, int lda_row) {
  for (int i = 0; i < rows; i++)
    for (int j = 0; j < cols; j++) a[(size_t)i * lda_row + j] = t[i + (size_t)j * rows];
}

static PyObject* py_qrf(PyObject* self, PyObject* args) {
  (void)self;
  int layout, m, n, lda;
  PyObject *oA, *otau;
  if (!PyArg_ParseTuple(args, "iiiOiO", &layout, &m, &n, &oA, &lda, &otau)) return NULL;
  buf_t A, tau;
  if (get_buf(oA, &A, "A")) return NULL;
  if (get_buf(otau, &tau, "tau")) { rel(&A); return NULL; }
  int out;
  double* a = (double*)A.view.buf;
  Py_BEGIN_ALLOW_THREADS
  if (layout == 102) {
    out = qrdm_dgeqrf(m, n, a, lda, (double*)tau.view.buf);
  } else if (layout == 101) {
    if (lda < (n > 1 ? n : 1)) out = -5;
    else {
      double* t = to_colmajor(a, m, n, lda);
      out = t ? qrdm_dgeqrf(m, n, t, m > 1 ? m : 1, (double*)tau.view.buf) : -1011;
      if (t && out == 0) from_colmajor(a, t, m, n, lda);
      free(t);
    }
  } else {
    out = -1;
  }
  Py_END_ALLOW_THREADS
  rel(&A); rel(&tau);
  return Py_BuildValue("i", out);
}

static PyObject* py_dormqr(PyObject* self, PyObject* args) {
  (void)self;
  int layout, m, n, k, lda, ldc;
  PyObject *oA, *otau, *oC;
  if (!PyArg_ParseTuple(args, "iiiiOiOOi", &layout, &m, &n, &k, &oA, &lda, &otau, &oC, &ldc))
    return NULL;
  buf_t A, tau, C;
  if (get_in_buf(oA, &A, "A")) return NULL;
  if (get_in_buf(otau, &tau, "tau")) { rel(&A); return NULL; }
  if (get_buf(oC, &C, "C")) { rel(&A); rel(&tau); return NULL; }
  int out;
  double* a = (double*)A.view.buf;
  double* c = (double*)C.view.buf;
  Py_BEGIN_ALLOW_THREADS
  if (layout == 102) {
    out = qrdm_dormqr(0, m, n, k, a, lda, (double*)tau.view.buf, c, ldc);
  } else if (layout == 101) {
    if (lda < (k > 1 ? k : 1)) out = -7;
    else if (ldc < (n > 1 ? n : 1)) out = -11;
    else {
      double* ta = to_colmajor(a, m, k, lda);
      double* tc = to_colmajor(c, m, n, ldc);
      out = (ta && tc) ? qrdm_dormqr(0, m, n, k, ta, m > 1 ? m : 1, (double*)tau.view.buf, tc,
                                     m > 1 ? m : 1)
                       : -1011;
      if (ta && tc && out == 0) from_colmajor(c, tc, m, n, ldc);
      free(ta);
      free(tc);
    }
  } else {
    out = -1;
  }
  Py_END_ALLOW_THREADS
  rel(&A); rel(&tau); rel(&C);
  return Py_BuildValue("i", out);
}

typedef int (*dgeqp3_fn)(int, int, int, double*, int, int*, double*);
static dgeqp3_fn g_dgeqp3 = NULL;

/* Locates LAPACKE_dgeqp3 in the OpenBLAS that ships with scipy. */
static int find_dgeqp3(void) {
  if (g_dgeqp3) return 0;
  PyObject* mod = PyImport_ImportModule("scipy");
  if (!mod) return -1;
  PyObject* file = PyObject_GetAttrString(mod, "__file__");
  Py_DECREF(mod);
  if (!file) return -1;
  const char* f = PyUnicode_AsUTF8(file);
  char dir[4096];
  snprintf(dir, sizeof(dir), "%s", f ? f : "");
  Py_DECREF(file);
  char* slash = strrchr(dir, '/');
  if (slash) *slash = 0;
  PyObject* os = PyImport_ImportModule("glob");
  if (!os) return -1;
  char pat[4200];
  snprintf(pat, sizeof(pat), "%s/../scipy.libs/libscipy_openblas*.so", dir);
  PyObject* lst = PyObject_CallMethod(os, "glob", "s", pat);
  Py_DECREF(os);
  if (!lst || !PyList_Check(lst) || PyList_Size(lst) < 1) {
    Py_XDECREF(lst);
    PyErr_SetString(PyExc_RuntimeError, "QRDM.QP3: host LAPACK (scipy OpenBLAS) not found");
    return -1;
  }
  const char* path = PyUnicode_AsUTF8(PyList_GetItem(lst, 0));
  void* h = dlopen(path, RTLD_LAZY | RTLD_LOCAL);
  Py_DECREF(lst);
  if (!h) {
    PyErr_Format(PyExc_RuntimeError, "QRDM.QP3: %s", dlerror());
    return -1;
  }
  g_dgeqp3 = (dgeqp3_fn)dlsym(h, "scipy_LAPACKE_dgeqp3");
  if (!g_dgeqp3) {
    PyErr_SetString(PyExc_RuntimeError, "QRDM.QP3: LAPACKE_dgeqp3 not exported");
    return -1;
  }
  return 0;
}

static PyObject* py_qp3(PyObject* self, PyObject* args) {
  (void)self;
  int layout, m, n, lda;
  PyObject *oA, *ojp, *otau;
  if (!PyArg_ParseTuple(args, "iiiOiOO", &layout, &m, &n, &oA, &lda, &ojp, &otau)) return NULL;
  if (find_dgeqp3()) return NULL;
  buf_t A, jp, tau;
  if (get_buf(oA, &A, "A")) return NULL;
  if (get_buf(ojp, &jp, "jpvt")) { rel(&A); return NULL; }
  if (get_buf(otau, &tau, "tau")) { rel(&A); rel(&jp); return NULL; }
  int out;
  Py_BEGIN_ALLOW_THREADS
  out = g_dgeqp3(layout, m, n, (double*)A.view.buf, lda, (int*)jp.view.buf,
                 (double*)tau.view.buf);
  Py_END_ALLOW_THREADS
  rel(&A); rel(&jp); rel(&tau);
  return Py_BuildValue("i", out);
}

static PyMethodDef methods[] = {
    {"init", py_init, METH_VARARGS, "initialize module"},
    {"QRF", py_qrf, METH_VARARGS, "QR without column pivoting (GPU)"},
    {"QP3", py_qp3, METH_VARARGS, "LAPACK QR with column pivoting (host LAPACKE_dgeqp3)"},
    {"QRDM", py_qrdm, METH_VARARGS, "QR with Deviation Maximization pivoting (GPU)"},
    {"DORMQR", py_dormqr, METH_VARARGS, "apply Q from the left, dormqr('L','N') (GPU)"},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef moddef = {PyModuleDef_HEAD_INIT, "QRDM",
                                    "QR with Deviation Maximization pivoting on MI355X.", -1,
                                    methods};

PyMODINIT_FUNC PyInit_QRDM(void) { return PyModule_Create(&moddef); }
