/*
 * grayhip_grid.hip -- resident numerical spacetimes (include/grayhip.h):
 * two device slots per spacetime, each holding one time level as
 * interleaved per-point Christoffels [p][40] and metric [p][10], and the
 * density [p].  Replaces
 * the 2 x 51 OpenCL images of sim-org/gray.h:92-93 and the upload half of
 * sim-org/io.c:302-533 (the HDF5 read is host code in libgray).
 */
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "launch.h"
#include "spacetime.h"

using namespace gray;

static int sfail(grayhip_spacetime *s, int code, const char *fmt, ...)
{
	static thread_local char g[256];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(s ? s->err : g, 256, fmt, ap);
	va_end(ap);
	return code;
}

extern "C" const char *grayhip_spacetime_error(const grayhip_spacetime *st)
{
	return st ? st->err : "spacetime: null handle or creation failed";
}

extern "C" void grayhip_spacetime_destroy(grayhip_spacetime *st)
{
	if (!st)
		return;
	(void)hipSetDevice(st->desc.device);
	for (int l = 0; l < 2; ++l) {
		(void)hipFree(st->gam[l]);
		(void)hipFree(st->met[l]);
		(void)hipFree(st->rho[l]);
	}
	(void)hipFree(st->d_tmp);
	free(st);
}

extern "C" int grayhip_spacetime_create(const grayhip_grid_desc *d, grayhip_spacetime **out)
{
	if (!d || !out)
		return GRAYHIP_EINVAL;
	*out = nullptr;
	for (int a = 0; a < 3; ++a)
		if (d->n[a] < 2 || !(d->hi[a] > d->lo[a]))
			return GRAYHIP_EINVAL;
	if (d->dtype != 4 && d->dtype != 8)
		return GRAYHIP_EINVAL;
	if (d->map != GRAYHIP_MAP_UNIFORM && d->map != GRAYHIP_MAP_FISHEYE)
		return GRAYHIP_EINVAL;
	const size_t np = (size_t)d->n[0] * d->n[1] * d->n[2];
	if (np > 0xffffffffu / 40)
		return GRAYHIP_EINVAL; /* corner indices are 32-bit */
	grayhip_spacetime *s = (grayhip_spacetime *)calloc(1, sizeof(grayhip_spacetime));
	if (!s)
		return GRAYHIP_ENOMEM;
	s->desc = *d;
	s->n_points = np;
	if (hipSetDevice(d->device) != hipSuccess) {
		free(s);
		return GRAYHIP_EDEVICE;
	}
	for (int l = 0; l < 2; ++l) {
		if (hipMalloc(&s->gam[l], np * 40 * d->dtype) != hipSuccess ||
		    hipMalloc(&s->met[l], np * 10 * d->dtype) != hipSuccess ||
		    hipMalloc(&s->rho[l], np * d->dtype) != hipSuccess) {
			grayhip_spacetime_destroy(s);
			return GRAYHIP_ENOMEM;
		}
	}
	s->frozen = 1;
	*out = s;
	return GRAYHIP_OK;
}

extern "C" int grayhip_spacetime_upload(grayhip_spacetime *s, int slot, double time,
                                        const void *gam, const void *met)
{
	if (!s || slot < 0 || slot > 1 || !gam || !met)
		return sfail(s, GRAYHIP_EINVAL, "upload: bad argument");
	const size_t sz = (size_t)s->desc.dtype;
	hipError_t e;
	if ((e = hipSetDevice(s->desc.device)) != hipSuccess ||
	    (e = hipMemcpy(s->gam[slot], gam, s->n_points * 40 * sz, hipMemcpyHostToDevice)) != hipSuccess ||
	    (e = hipMemcpy(s->met[slot], met, s->n_points * 10 * sz, hipMemcpyHostToDevice)) != hipSuccess)
		return sfail(s, GRAYHIP_EDEVICE, "upload: %s", hipGetErrorString(e));
	s->time[slot] = time;
	s->loaded[slot] = 1;
	s->has_rho[slot] = 0;
	if (slot == 1)
		s->frozen = 0;
	return GRAYHIP_OK;
}

extern "C" int grayhip_spacetime_upload_rho(grayhip_spacetime *s, int slot, const void *rho)
{
	if (!s || slot < 0 || slot > 1 || !rho)
		return sfail(s, GRAYHIP_EINVAL, "upload_rho: bad argument");
	if (!s->loaded[slot])
		return sfail(s, GRAYHIP_ESTATE, "upload_rho: slot %d holds no level", slot);
	hipError_t e;
	if ((e = hipSetDevice(s->desc.device)) != hipSuccess ||
	    (e = hipMemcpy(s->rho[slot], rho, s->n_points * (size_t)s->desc.dtype,
	                   hipMemcpyHostToDevice)) != hipSuccess)
		return sfail(s, GRAYHIP_EDEVICE, "upload_rho: %s", hipGetErrorString(e));
	s->has_rho[slot] = 1;
	return GRAYHIP_OK;
}

extern "C" int grayhip_spacetime_shift(grayhip_spacetime *s, int to_t1)
{
	if (!s)
		return GRAYHIP_EINVAL;
	/* swap the slots; the one that is not current any more is free */
	void *g = s->gam[0], *m = s->met[0], *r = s->rho[0];
	double t = s->time[0];
	int l = s->loaded[0], hr = s->has_rho[0];
	s->gam[0] = s->gam[1]; s->met[0] = s->met[1]; s->rho[0] = s->rho[1];
	s->time[0] = s->time[1]; s->loaded[0] = s->loaded[1]; s->has_rho[0] = s->has_rho[1];
	s->gam[1] = g; s->met[1] = m; s->rho[1] = r; s->time[1] = t; s->loaded[1] = l;
	s->has_rho[1] = hr;
	(void)to_t1; /* both directions are a swap; the caller reloads the other slot */
	return GRAYHIP_OK;
}

extern "C" int grayhip_spacetime_set_frozen(grayhip_spacetime *s, int frozen)
{
	if (!s)
		return GRAYHIP_EINVAL;
	if (!frozen && !(s->loaded[0] && s->loaded[1]))
		return sfail(s, GRAYHIP_ESTATE, "two levels needed for time interpolation");
	s->frozen = frozen != 0;
	return GRAYHIP_OK;
}

extern "C" size_t grayhip_spacetime_bytes(const grayhip_spacetime *s)
{
	return s ? 2 * s->n_points * 51 * (size_t)s->desc.dtype : 0;
}

int grayhip_spacetime_has_rho(const grayhip_spacetime *s)
{
	return s->has_rho[0] && (s->frozen || s->has_rho[1]);
}

GridDesc spacetime_view(const grayhip_spacetime *s, double t_off, int raw_icond, int ref_rule)
{
	GridDesc g;
	memset(&g, 0, sizeof(g));
	g.dtype = s->desc.dtype;
	for (int l = 0; l < 2; ++l) {
		g.gam[l] = s->gam[s->frozen ? 0 : l];
		g.met[l] = s->met[s->frozen ? 0 : l];
		g.rho[l] = s->rho[s->frozen ? 0 : l];
	}
	g.map = s->desc.map;
	for (int a = 0; a < 3; ++a) {
		g.n[a] = s->desc.n[a];
		g.lo[a] = s->desc.lo[a];
		g.hi[a] = s->desc.hi[a];
		g.dx[a] = (g.hi[a] - g.lo[a]) / (double)(g.n[a] - 1);
		g.idx[a] = (double)(g.n[a] - 1) / (g.hi[a] - g.lo[a]);
		g.fb[a] = cbrt(asinh(g.lo[a]));
		g.fa[a] = (cbrt(asinh(g.hi[a])) - g.fb[a]) / (double)(g.n[a] - 1);
	}
	g.t1 = s->time[0];
	g.t2 = s->frozen ? s->time[0] : s->time[1];
	g.t_off = t_off;
	g.raw_icond = raw_icond;
	g.ref_rule = ref_rule;
	return g;
}

extern "C" int grayhip_spacetime_sample(grayhip_spacetime *s, int n, const double *events,
                                        int which, int comp, double t_init, double *out)
{
	if (!s || n < 0 || !events || !out || which < 0 || which > 2 || comp < 0 ||
	    comp >= (which == 0 ? 40 : (which == 1 ? 10 : 1)))
		return sfail(s, GRAYHIP_EINVAL, "sample: bad argument");
	if (!s->loaded[0])
		return sfail(s, GRAYHIP_ESTATE, "sample: no level loaded");
	if (which == 2 && !grayhip_spacetime_has_rho(s))
		return sfail(s, GRAYHIP_ESTATE, "sample: no density uploaded");
	hipError_t e;
	if ((e = hipSetDevice(s->desc.device)) != hipSuccess)
		return sfail(s, GRAYHIP_EDEVICE, "sample: %s", hipGetErrorString(e));
	size_t need = (size_t)n * 5;
	if (need > s->tmp_cap) {
		(void)hipFree(s->d_tmp);
		s->d_tmp = nullptr;
		s->tmp_cap = 0;
		if (hipMalloc(&s->d_tmp, need * sizeof(double)) != hipSuccess)
			return sfail(s, GRAYHIP_ENOMEM, "sample: allocation failed");
		s->tmp_cap = need;
	}
	GridDesc g = spacetime_view(s, t_init, 0, 0);
	if ((e = hipMemcpy(s->d_tmp, events, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
	    (e = launch_grid_sample_f64(g, s->d_tmp, which, comp, s->d_tmp + 4 * (size_t)n, n, 0)) != hipSuccess ||
	    (e = hipMemcpy(out, s->d_tmp + 4 * (size_t)n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
		return sfail(s, GRAYHIP_EDEVICE, "sample: %s", hipGetErrorString(e));
	return GRAYHIP_OK;
}
