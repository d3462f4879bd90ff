/* devmem.c -- process-wide cache of device buffers.
 *
 * Operators, their Davidson subspaces, slot buffers and solver scratch are
 * allocated through dh_dmalloc and released through dh_dfree.  A released
 * buffer is kept (up to DH_CACHE_MAX bytes in all) and handed to the next
 * request it fits (at most twice the requested size), so that a sequence of
 * solves -- one request after another -- does not pay the driver's
 * allocation and release of gigabytes per solve.  The caller synchronises
 * the stream that last used a buffer before releasing it.  dh_dcache_trim
 * returns every cached buffer to the driver (before a large reservation,
 * and when an allocation fails). */
#include <pthread.h>
#include <stdlib.h>

#include "internal.h"

#define DH_CACHE_MAX ((size_t)64 << 30)

typedef struct dblk {
    void *p;
    size_t size;
    struct dblk *next;
} dblk;

static dblk *g_live, *g_free;
static size_t g_free_bytes;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

/* sizes rounded so that buffers of nearby sizes are interchangeable */
static size_t round_size(size_t bytes)
{
    const size_t g = bytes >= ((size_t)1 << 20) ? ((size_t)1 << 20) : 4096;
    return (bytes + g - 1) / g * g;
}

static void trim_locked(void)
{
    while (g_free) {
        dblk *b = g_free;
        g_free = b->next;
        hipFree(b->p);
        free(b);
    }
    g_free_bytes = 0;
}

void dh_dcache_trim(void)
{
    pthread_mutex_lock(&g_lock);
    trim_locked();
    pthread_mutex_unlock(&g_lock);
}

hipError_t dh_dmalloc(void **p, size_t bytes)
{
    *p = NULL;
    const size_t want = round_size(bytes ? bytes : 1);
    pthread_mutex_lock(&g_lock);
    dblk **best = NULL;
    for (dblk **q = &g_free; *q; q = &(*q)->next)
        if ((*q)->size >= want && (*q)->size <= 2 * want && (!best || (*q)->size < (*best)->size))
            best = q;
    if (best) {
        dblk *b = *best;
        *best = b->next;
        g_free_bytes -= b->size;
        b->next = g_live;
        g_live = b;
        *p = b->p;
        pthread_mutex_unlock(&g_lock);
        return hipSuccess;
    }
    dblk *b = malloc(sizeof *b);
    if (!b) {
        pthread_mutex_unlock(&g_lock);
        return hipErrorOutOfMemory;
    }
    hipError_t e = hipMalloc(&b->p, want);
    if (e != hipSuccess) { /* give the cache back to the driver and try once more */
        (void)hipGetLastError();
        trim_locked();
        e = hipMalloc(&b->p, want);
    }
    if (e != hipSuccess) {
        free(b);
        pthread_mutex_unlock(&g_lock);
        return e;
    }
    b->size = want;
    b->next = g_live;
    g_live = b;
    *p = b->p;
    pthread_mutex_unlock(&g_lock);
    return hipSuccess;
}

void dh_dfree(void *p)
{
    if (!p) return;
    pthread_mutex_lock(&g_lock);
    for (dblk **q = &g_live; *q; q = &(*q)->next) {
        if ((*q)->p != p) continue;
        dblk *b = *q;
        *q = b->next;
        if (g_free_bytes + b->size <= DH_CACHE_MAX) {
            b->next = g_free;
            g_free = b;
            g_free_bytes += b->size;
        } else {
            hipFree(b->p);
            free(b);
        }
        pthread_mutex_unlock(&g_lock);
        return;
    }
    pthread_mutex_unlock(&g_lock);
    hipFree(p); /* not from dh_dmalloc */
}
