/*
 * _rowsmap: the triangle map {tid: (x, y, id_a, id_b, id_c)} of Python ints
 * built in C from a (T, 5) int32 row buffer -- the dictionary the reference's
 * binding returns (src/cuda/bindings/binding_helpers.hpp:61-87).  Building it
 * with Python-level loops costs about 1 us per triangle (64 ms of a 65 ms
 * triangulate() at 1510 x 1018 with 30,000 seeds).  Ints are immutable, so
 * equal values share one object (a seed id sits in about six triangles), and
 * the cycle collector, which would rescan the growing dictionary's tuples
 * again and again, is paused while the map is built.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static PyObject* rows_to_map(PyObject* self, PyObject* arg)
{
    (void)self;
    Py_buffer view;
    if (PyObject_GetBuffer(arg, &view, PyBUF_C_CONTIGUOUS | PyBUF_FORMAT) != 0) return NULL;
    if (view.itemsize != 4 || view.len % 20 != 0 ||
        (view.format && strcmp(view.format, "i") != 0 && strcmp(view.format, "<i") != 0 &&
         strcmp(view.format, "=i") != 0)) {
        PyBuffer_Release(&view);
        PyErr_SetString(PyExc_ValueError, "rows must be a C-contiguous (T, 5) int32 buffer");
        return NULL;
    }
    const int32_t* rows = (const int32_t*)view.buf;
    const Py_ssize_t n = view.len / 20;
    /* One shared object per value in [lo, hi] when that range is not much
     * larger than the number of values. */
    int64_t lo = 0, hi = -1;
    for (Py_ssize_t i = 0; i < 5 * n; ++i) {
        if (i == 0 || rows[i] < lo) lo = rows[i];
        if (i == 0 || rows[i] > hi) hi = rows[i];
    }
    PyObject** cache = NULL;
    if (n > 0 && hi - lo < 4 * (int64_t)n + 65536)
        cache = (PyObject**)calloc((size_t)(hi - lo + 1), sizeof(PyObject*));
    PyObject* d = _PyDict_NewPresized(n);
    const int gc_was_on = PyGC_Disable();
    int ok = d != NULL;
    const int32_t* r = rows;
    for (Py_ssize_t i = 0; ok && i < n; ++i, r += 5) {
        PyObject* t = PyTuple_New(5);
        PyObject* k = t ? PyLong_FromSsize_t(i) : NULL;
        ok = k != NULL;
        for (int j = 0; ok && j < 5; ++j) {
            PyObject* v;
            if (cache) {
                PyObject** slot = &cache[r[j] - lo];
                if (!*slot) *slot = PyLong_FromLong(r[j]);
                v = *slot;
                Py_XINCREF(v);
            } else {
                v = PyLong_FromLong(r[j]);
            }
            if (!v) ok = 0;
            else PyTuple_SET_ITEM(t, j, v);   /* steals v */
        }
        if (ok) ok = PyDict_SetItem(d, k, t) == 0;
        Py_XDECREF(k);
        Py_XDECREF(t);
    }
    if (gc_was_on) PyGC_Enable();
    if (cache) {
        for (int64_t v = 0; v <= hi - lo; ++v) Py_XDECREF(cache[v]);
        free(cache);
    }
    PyBuffer_Release(&view);
    if (!ok) {
        Py_XDECREF(d);
        if (!PyErr_Occurred()) PyErr_NoMemory();
        return NULL;
    }
    return d;
}

static PyMethodDef methods[] = {
    {"rows_to_map", rows_to_map, METH_O,
     "rows_to_map(rows) -> {tid: (x, y, id_a, id_b, id_c)} from a (T, 5) int32 buffer"},
    {NULL, NULL, 0, NULL},
};

static struct PyModuleDef module = {PyModuleDef_HEAD_INIT, "_rowsmap", NULL, -1, methods,
                                    NULL, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit__rowsmap(void) { return PyModule_Create(&module); }
