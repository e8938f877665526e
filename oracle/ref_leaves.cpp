// Driver exporting the reference's own leaf functions (compiled from the
// reference sources with -DHOST) with C linkage, so the oracle's restatement
// can be checked against them. Test infrastructure only.
#include <random.hh>
#include <hash.hh>
#include <algebra/complex.hh>
#include <algebra/quaternion.hh>

extern "C" {

unsigned ref_rng_next(unsigned *state) {
    Rng r = rng_init(*state);
    unsigned v = rng_next(&r);
    *state = r.state;
    return v;
}

float ref_random_uniform(unsigned *state) {
    Rng r = rng_init(*state);
    float u = random_uniform(&r);
    *state = r.state;
    return u;
}

void ref_random_hemisphere_cosine(unsigned *state, float *out3) {
    Rng r = rng_init(*state);
    real3 v = random_hemisphere_cosine(&r);
    *state = r.state;
    out3[0] = v.x;
    out3[1] = v.y;
    out3[2] = v.z;
}

unsigned ref_hash(const unsigned *xs, int n) {
    Hasher h = hash_init();
    for (int i = 0; i < n; ++i) hash_uint(&h, xs[i]);
    return hash_finish(&h);
}

void ref_q_mul(const float *a, const float *b, float *out) {
    quat r = q_mul(make_real4(a[0], a[1], a[2], a[3]), make_real4(b[0], b[1], b[2], b[3]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

void ref_q_div(const float *a, const float *b, float *out) {
    quat r = q_div(make_real4(a[0], a[1], a[2], a[3]), make_real4(b[0], b[1], b[2], b[3]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

void ref_c_mul(const float *a, const float *b, float *out) {
    comp r = c_mul(make_real2(a[0], a[1]), make_real2(b[0], b[1]));
    out[0] = r.x;
    out[1] = r.y;
}

}  // extern "C"
