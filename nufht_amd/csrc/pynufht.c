/*
 * _pynufht: the reference's Python entry point (src/_pynufht.c:246-271) on
 * the MI355X engine.  Same module name, functions, keyword arguments,
 * array coercions, return shapes and error behaviour:
 *
 *   nufht(nu, rs, cs, ws, tol=1e-8)            -> float64 (n,)
 *   nufht_batch_loop(nu, rs, cs, ws, tol=1e-8) -> float64 (n, batch), one
 *       fresh plan per column (the reference's "plain" path, :65-142)
 *   nufht_batch(nu, rs, cs, ws, tol=1e-8)      -> float64 (n, batch), one
 *       plan for all columns (:144-244)
 *
 * As in the reference, nufht and nufht_batch_loop go through nufht(), which
 * reports plan failures on stderr and returns zeros; nufht_batch raises
 * RuntimeError.  ValueError for a cs that is not 2-D or whose row count is
 * not len(rs).  Unlike the reference the GIL is released while the device
 * works.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>
#include <stdlib.h>
#include <string.h>

#include "fast_hankel_nufht.h"
#include "nufht_amd.h"

static char *kwlist[] = {"nu", "rs", "cs", "ws", "tol", NULL};

static int parse(PyObject *args, PyObject *kw, double *nu, PyArrayObject **rs, PyObject **cs_obj,
                 PyArrayObject **ws, double *tol)
{
    PyObject *r = NULL, *w = NULL;
    *tol = 1e-8;
    if (!PyArg_ParseTupleAndKeywords(args, kw, "dOOO|d", kwlist, nu, &r, cs_obj, &w, tol)) return -1;
    *rs = (PyArrayObject *)PyArray_FROM_OTF(r, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    *ws = (PyArrayObject *)PyArray_FROM_OTF(w, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    if (!*rs || !*ws) {
        Py_XDECREF(*rs);
        Py_XDECREF(*ws);
        return -1;
    }
    return 0;
}

static PyObject *py_nufht(PyObject *self, PyObject *args, PyObject *kw)
{
    (void)self;
    double nu, tol;
    PyArrayObject *rs, *ws;
    PyObject *cs_obj;
    if (parse(args, kw, &nu, &rs, &cs_obj, &ws, &tol)) return NULL;
    PyArrayObject *cs = (PyArrayObject *)PyArray_FROM_OTF(cs_obj, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    if (!cs) { Py_DECREF(rs); Py_DECREF(ws); return NULL; }
    int m = (int)PyArray_DIM(rs, 0), n = (int)PyArray_DIM(ws, 0);
    NufhtOptions opt = nufht_default_options();
    opt.tol = tol;
    double *out;
    Py_BEGIN_ALLOW_THREADS
    out = nufht(nu, (const double *)PyArray_DATA(rs), m, (const double *)PyArray_DATA(cs),
                (const double *)PyArray_DATA(ws), n, &opt);
    Py_END_ALLOW_THREADS
    if (!out) {
        Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
        PyErr_SetString(PyExc_RuntimeError, "nufht returned NULL");
        return NULL;
    }
    npy_intp dims[1] = {n};
    PyObject *res = PyArray_SimpleNew(1, dims, NPY_DOUBLE);
    if (res) memcpy(PyArray_DATA((PyArrayObject *)res), out, sizeof(double) * (size_t)n);
    free(out);
    Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
    return res ? res : PyErr_NoMemory();
}

static int check_cs(PyArrayObject *cs, int m, int *batch)
{
    if (PyArray_NDIM(cs) != 2) {
        PyErr_SetString(PyExc_ValueError, "cs must be 2-D array with shape (m, batch)");
        return -1;
    }
    if ((int)PyArray_DIM(cs, 0) != m) {
        PyErr_SetString(PyExc_ValueError, "cs first dimension must match rs length");
        return -1;
    }
    *batch = (int)PyArray_DIM(cs, 1);
    return 0;
}

static PyObject *py_nufht_batch_loop(PyObject *self, PyObject *args, PyObject *kw)
{
    (void)self;
    double nu, tol;
    PyArrayObject *rs, *ws;
    PyObject *cs_obj;
    if (parse(args, kw, &nu, &rs, &cs_obj, &ws, &tol)) return NULL;
    PyArrayObject *cs = (PyArrayObject *)PyArray_FROM_OTF(cs_obj, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    if (!cs) { Py_DECREF(rs); Py_DECREF(ws); return NULL; }
    int m = (int)PyArray_DIM(rs, 0), n = (int)PyArray_DIM(ws, 0), batch;
    if (check_cs(cs, m, &batch)) { Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs); return NULL; }
    npy_intp dims[2] = {n, batch};
    PyObject *res = PyArray_SimpleNew(2, dims, NPY_DOUBLE);
    double *col = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    if (!res || !col) {
        Py_XDECREF(res); free(col);
        Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
        return PyErr_NoMemory();
    }
    NufhtOptions opt = nufht_default_options();
    opt.tol = tol;
    const double *c = (const double *)PyArray_DATA(cs);
    double *g = (double *)PyArray_DATA((PyArrayObject *)res);
    int failed = 0;
    Py_BEGIN_ALLOW_THREADS
    for (int b = 0; b < batch && !failed; ++b) {
        for (int i = 0; i < m; ++i) col[i] = c[(size_t)i * batch + b];
        double *out = nufht(nu, (const double *)PyArray_DATA(rs), m, col,
                            (const double *)PyArray_DATA(ws), n, &opt);
        if (!out) { failed = 1; break; }
        for (int i = 0; i < n; ++i) g[(size_t)i * batch + b] = out[i];
        free(out);
    }
    Py_END_ALLOW_THREADS
    free(col);
    Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
    if (failed) {
        Py_DECREF(res);
        PyErr_SetString(PyExc_RuntimeError, "nufht returned NULL inside batch loop");
        return NULL;
    }
    return res;
}

static PyObject *py_nufht_batch(PyObject *self, PyObject *args, PyObject *kw)
{
    (void)self;
    double nu, tol;
    PyArrayObject *rs, *ws;
    PyObject *cs_obj;
    if (parse(args, kw, &nu, &rs, &cs_obj, &ws, &tol)) return NULL;
    PyArrayObject *cs = (PyArrayObject *)PyArray_FROM_OTF(
        cs_obj, NPY_DOUBLE, NPY_ARRAY_F_CONTIGUOUS | NPY_ARRAY_ALIGNED);
    if (!cs) { Py_DECREF(rs); Py_DECREF(ws); return NULL; }
    int m = (int)PyArray_DIM(rs, 0), n = (int)PyArray_DIM(ws, 0), batch;
    if (check_cs(cs, m, &batch)) { Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs); return NULL; }
    NufhtOptions opt = nufht_default_options();
    opt.tol = tol;
    NufhtPlan plan;
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = nufht_plan_init(&plan, nu, (const double *)PyArray_DATA(rs), m,
                         (const double *)PyArray_DATA(ws), n, &opt);
    Py_END_ALLOW_THREADS
    if (rc != 0) {
        Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
        PyErr_SetString(PyExc_RuntimeError, "nufht_plan_init failed");
        return NULL;
    }
    NufhtScratch scratch;
    if (nufht_scratch_init(&scratch, &plan) != 0) {
        nufht_plan_free(&plan);
        Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
        PyErr_SetString(PyExc_RuntimeError, "nufht_scratch_init failed");
        return NULL;
    }
    double *gbuf = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(batch > 0 ? batch : 1));
    if (!gbuf) {
        nufht_scratch_free(&scratch);
        nufht_plan_free(&plan);
        Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
        return PyErr_NoMemory();
    }
    Py_BEGIN_ALLOW_THREADS
    rc = nufht_batch(&plan, (const double *)PyArray_DATA(cs), m, gbuf, n, batch, &scratch);
    Py_END_ALLOW_THREADS
    if (rc != 0) {
        free(gbuf);
        nufht_scratch_free(&scratch);
        nufht_plan_free(&plan);
        Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
        PyErr_Format(PyExc_RuntimeError, "nufht_batch returned %d (%s)", rc, nufht_amd_last_error());
        return NULL;
    }
    npy_intp dims[2] = {n, batch};
    PyObject *res = PyArray_SimpleNew(2, dims, NPY_DOUBLE);
    if (res) {
        double *g = (double *)PyArray_DATA((PyArrayObject *)res);
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < n; ++i) g[(size_t)i * batch + b] = gbuf[(size_t)b * n + i];
    }
    free(gbuf);
    nufht_scratch_free(&scratch);
    nufht_plan_free(&plan);
    Py_DECREF(rs); Py_DECREF(ws); Py_DECREF(cs);
    return res ? res : PyErr_NoMemory();
}

static PyMethodDef methods[] = {
    {"nufht", (PyCFunction)(void (*)(void))py_nufht, METH_VARARGS | METH_KEYWORDS,
     "nufht(nu, rs, cs, ws, tol=1e-8) -> array(n)"},
    {"nufht_batch_loop", (PyCFunction)(void (*)(void))py_nufht_batch_loop,
     METH_VARARGS | METH_KEYWORDS,
     "nufht_batch_loop(nu, rs, cs_matrix, ws, tol=1e-8) -> array(n, batch)"},
    {"nufht_batch", (PyCFunction)(void (*)(void))py_nufht_batch, METH_VARARGS | METH_KEYWORDS,
     "nufht_batch(nu, rs, cs_matrix, ws, tol=1e-8) -> array(n, batch)"},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef moddef = {PyModuleDef_HEAD_INIT, "_pynufht",
                                    "NUFHT on AMD Instinct MI355X (drop-in for _pynufht)", -1,
                                    methods, NULL, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit__pynufht(void)
{
    PyObject *m = PyModule_Create(&moddef);
    if (!m) return NULL;
    import_array();
    return m;
}
