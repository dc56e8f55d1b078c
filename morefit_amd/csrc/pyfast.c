/*
 * _fast: the per-call evaluation of morefit_amd.Batch.eval as a CPython extension.
 *
 * A ctypes call with numpy argument handling costs ~3 us per call, about a tenth of a
 * one-fit -2lnL + gradient evaluation of 10^7 events on MI355X (~24 us in C++). This module
 * makes the same call through the C API. It links nothing: morefit_amd hands it the address of
 * mfh_batch_eval (include/morefit_capi.h) from the library it has loaded.
 *
 *   _fast.init(address_of_mfh_batch_eval)
 *   _fast.eval(batch_handle, kind, theta, use_graph, nseg, ntheta, nout) -> ndarray (nseg, nout)
 *                                                                      or int (nonzero status)
 *
 * theta: anything numpy converts to float64 with nseg * ntheta values, or ntheta values that
 * every segment shares.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>
#include <string.h>

typedef int (*batch_eval_fn)(void* batch, int kind, const double* theta, int use_graph, double* out);
static batch_eval_fn g_eval = NULL;

static PyObject* fast_init(PyObject* self, PyObject* args) {
  unsigned long long addr = 0;
  if (!PyArg_ParseTuple(args, "K", &addr)) return NULL;
  g_eval = (batch_eval_fn)(uintptr_t)addr;
  Py_RETURN_NONE;
}

static PyObject* fast_eval(PyObject* self, PyObject* const* args, Py_ssize_t nargs) {
  if (nargs != 7) {
    PyErr_SetString(PyExc_TypeError, "eval(batch, kind, theta, use_graph, nseg, ntheta, nout)");
    return NULL;
  }
  if (!g_eval) {
    PyErr_SetString(PyExc_RuntimeError, "_fast.init was not called");
    return NULL;
  }
  void* batch = PyLong_AsVoidPtr(args[0]);
  const long kind = PyLong_AsLong(args[1]);
  const long use_graph = PyLong_AsLong(args[3]);
  const long nseg = PyLong_AsLong(args[4]), ntheta = PyLong_AsLong(args[5]), nout = PyLong_AsLong(args[6]);
  if (PyErr_Occurred()) return NULL;
  if (!batch || nseg < 0 || ntheta < 0 || nout < 0) {
    PyErr_SetString(PyExc_ValueError, "eval: bad batch handle or sizes");
    return NULL;
  }
  PyArrayObject* th = (PyArrayObject*)PyArray_FROMANY(args[2], NPY_DOUBLE, 0, 2, NPY_ARRAY_C_CONTIGUOUS | NPY_ARRAY_ALIGNED);
  if (!th) return NULL;
  const npy_intp size = PyArray_SIZE(th);
  const double* tp = (const double*)PyArray_DATA(th);
  double* shared = NULL;
  if (size != (npy_intp)nseg * ntheta) {
    if (size != ntheta) {
      Py_DECREF(th);
      PyErr_Format(PyExc_ValueError, "eval: theta has %zd values, not %ld x %ld or %ld", (Py_ssize_t)size, nseg, ntheta, ntheta);
      return NULL;
    }
    shared = (double*)PyMem_Malloc(sizeof(double) * (size_t)(nseg * ntheta > 0 ? nseg * ntheta : 1));
    if (!shared) {
      Py_DECREF(th);
      return PyErr_NoMemory();
    }
    for (long s = 0; s < nseg; s++) memcpy(shared + s * ntheta, tp, sizeof(double) * (size_t)ntheta);
    tp = shared;
  }
  npy_intp dims[2] = {nseg, nout};
  PyArrayObject* out = (PyArrayObject*)PyArray_SimpleNew(2, dims, NPY_DOUBLE);
  if (!out) {
    PyMem_Free(shared);
    Py_DECREF(th);
    return NULL;
  }
  int rc;
  Py_BEGIN_ALLOW_THREADS
  rc = g_eval(batch, (int)kind, tp, (int)use_graph, (double*)PyArray_DATA(out));
  Py_END_ALLOW_THREADS
  PyMem_Free(shared);
  Py_DECREF(th);
  if (rc) {
    Py_DECREF(out);
    return PyLong_FromLong(rc);
  }
  return (PyObject*)out;
}

static PyMethodDef fast_methods[] = {
    {"init", fast_init, METH_VARARGS, "init(address of mfh_batch_eval)"},
    {"eval", (PyCFunction)(void (*)(void))fast_eval, METH_FASTCALL,
     "eval(batch, kind, theta, use_graph, nseg, ntheta, nout) -> ndarray (nseg, nout), or the status when nonzero"},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef fast_module = {PyModuleDef_HEAD_INIT, "_fast", "per-call evaluation through the C API", -1, fast_methods};

PyMODINIT_FUNC PyInit__fast(void) {
  import_array();
  return PyModule_Create(&fast_module);
}
