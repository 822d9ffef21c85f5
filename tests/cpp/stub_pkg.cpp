// Test double of the C ABI for checking the C++ facade's host logic on a machine
// without a GPU: it keeps bodies in host memory and advances them ballistically
// (x += v dt after v += g dt for dynamic bodies), with no collisions and no events.
// Linked only into tests/cpp/facade_logic; never part of the library.
#include "physkit_amd.h"

#include <algorithm>
#include <string>
#include <vector>

struct pkg_ctx
{
    pkg_config cfg{};
    std::string err;
    uint32_t n_shapes = 0, n = 0;
    std::vector<uint8_t> st;
    std::vector<double> pos, vel, ang, quat;
    int uploads = 0;
    std::vector<uint32_t> joints; // live joint keys, ascending
};

extern "C"
{
pkg_status pkg_create(pkg_ctx **out, const pkg_config *cfg)
{
    if (!out || !cfg) return PKG_INVALID_ARG;
    *out = new pkg_ctx();
    (*out)->cfg = *cfg;
    return PKG_OK;
}
pkg_status pkg_destroy(pkg_ctx *ctx)
{
    delete ctx;
    return PKG_OK;
}
const char *pkg_last_error(const pkg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }
pkg_status pkg_upload_shapes(pkg_ctx *ctx, uint32_t n_shapes, const uint32_t *offsets, const double *verts)
{
    if (!ctx || !offsets || (n_shapes && !verts)) return PKG_INVALID_ARG;
    ctx->n_shapes = n_shapes;
    return PKG_OK;
}
pkg_status pkg_upload_bodies(pkg_ctx *ctx, uint32_t n, const pkg_body_desc *d)
{
    if (!ctx || !d) return PKG_INVALID_ARG;
    for (uint32_t i = 0; i < n; ++i)
        if (d->shape[i] >= ctx->n_shapes) return PKG_INVALID_ARG;
    ctx->n = n;
    ctx->st.assign(d->is_static, d->is_static + n);
    ctx->pos.assign(d->pos, d->pos + 3 * n);
    ctx->vel.assign(d->vel, d->vel + 3 * n);
    ctx->ang.assign(d->angvel, d->angvel + 3 * n);
    ctx->quat.assign(d->quat, d->quat + 4 * n);
    ctx->uploads++;
    return PKG_OK;
}
pkg_status pkg_set_body_state(pkg_ctx *, uint32_t, const uint32_t *, const double *, const double *,
                              const double *, const double *)
{ return PKG_OK; }
pkg_status pkg_step(pkg_ctx *ctx, double dt, int nsteps)
{
    if (!ctx) return PKG_INVALID_ARG;
    // dt = 1234.5 stands for a step that fails part-way (the library's PKG_OOM case)
    const bool fail = dt == 1234.5;
    for (int s = 0; s < nsteps; ++s)
        for (uint32_t i = 0; i < ctx->n; ++i)
        {
            if (ctx->st[i]) continue;
            for (int k = 0; k < 3; ++k)
            {
                ctx->vel[3 * i + k] += ctx->cfg.gravity[k] * dt;
                ctx->pos[3 * i + k] += ctx->vel[3 * i + k] * dt;
            }
        }
    return fail ? PKG_OOM : PKG_OK;
}
pkg_status pkg_download_bodies(pkg_ctx *ctx, const pkg_body_state *out)
{
    if (!ctx || !out) return PKG_INVALID_ARG;
    for (size_t i = 0; i < 3 * ctx->n; ++i)
    {
        if (out->pos) out->pos[i] = ctx->pos[i];
        if (out->vel) out->vel[i] = ctx->vel[i];
        if (out->angvel) out->angvel[i] = ctx->ang[i];
    }
    if (out->quat)
        for (size_t i = 0; i < 4 * ctx->n; ++i) out->quat[i] = ctx->quat[i];
    return PKG_OK;
}
pkg_status pkg_get_pairs(pkg_ctx *, uint64_t *, size_t, size_t *n)
{
    *n = 0;
    return PKG_OK;
}
pkg_status pkg_get_manifolds(pkg_ctx *, int32_t *, double *, size_t, size_t *n)
{
    *n = 0;
    return PKG_OK;
}
pkg_status pkg_get_events(pkg_ctx *, uint64_t *, int32_t *, size_t, size_t *n)
{
    *n = 0;
    return PKG_OK;
}
pkg_status pkg_add_joints(pkg_ctx *ctx, const pkg_joint_desc *d, size_t n)
{
    for (size_t i = 0; i < n; ++i)
    {
        uint32_t k = PKG_JOINT_KEY(d[i].type, d[i].slot);
        if (d[i].a >= ctx->n || d[i].b >= ctx->n || d[i].a == d[i].b ||
            std::binary_search(ctx->joints.begin(), ctx->joints.end(), k))
            return PKG_INVALID_ARG;
        ctx->joints.insert(std::lower_bound(ctx->joints.begin(), ctx->joints.end(), k), k);
    }
    return PKG_OK;
}
pkg_status pkg_remove_joints(pkg_ctx *ctx, const uint32_t *keys, size_t n)
{
    for (size_t i = 0; i < n; ++i)
    {
        auto it = std::lower_bound(ctx->joints.begin(), ctx->joints.end(), keys[i]);
        if (it == ctx->joints.end() || *it != keys[i]) return PKG_INVALID_ARG;
        ctx->joints.erase(it);
    }
    return PKG_OK;
}
pkg_status pkg_get_joints(pkg_ctx *ctx, uint32_t *keys, double *impulses, size_t cap, size_t *n)
{
    *n = ctx->joints.size();
    for (size_t i = 0; i < std::min(cap, ctx->joints.size()); ++i)
    {
        if (keys) keys[i] = ctx->joints[i];
        if (impulses)
            for (size_t r = 0; r < 6; ++r) impulses[6 * i + r] = 0.0;
    }
    return PKG_OK;
}
pkg_status pkg_raycast(pkg_ctx *, uint32_t n_rays, const double *, const double *, const double *,
                       const uint32_t *, uint32_t *offsets, uint32_t *, double *, size_t, size_t *total)
{
    for (uint32_t i = 0; i <= n_rays; ++i) offsets[i] = 0;
    *total = 0;
    return PKG_OK;
}
pkg_status pkg_query_aabb(pkg_ctx *, uint32_t n_boxes, const double *, const uint32_t *, uint32_t *offsets,
                          uint32_t *, size_t, size_t *total)
{
    for (uint32_t i = 0; i <= n_boxes; ++i) offsets[i] = 0;
    *total = 0;
    return PKG_OK;
}
pkg_status pkg_get_counters(pkg_ctx *ctx, pkg_counters *out)
{
    *out = pkg_counters{};
    out->pairs = static_cast<uint64_t>(ctx->uploads); // the facade test reads the upload count here
    return PKG_OK;
}
}
