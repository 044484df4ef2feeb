// Request loop with continuous batching over the CL-FTS pool and TILD
// engines (mft_server_* in mft.h).
//
// Admission is first-fit against a device-memory budget, shortest work
// first: queued requests are tried in order of their estimated device
// traffic (iterations or steps times the bytes one of them moves), and a
// request passed over by kAgeLimit admissions of others goes ahead of all
// requests that have not waited as long, so long requests are not starved.
// A request is admitted when the estimate of its device buffers fits beside
// the admitted work.  Admitted CL-FTS solves join the pool after their own
// first pass; the pool runs in passes of up to max_iters iterations, cut
// short so that no solve runs past its requested count, and solves that
// reach their count are read out and leave.  TILD runs are queued whole on
// their engine's stream when admitted and finish on their own; the pool and
// the engines share the device concurrently.
#include <algorithm>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.h"

namespace mft {

struct FtsBatch;
int fts_pool_create(FtsBatch** b, int nslots);
int fts_pool_destroy(FtsBatch* b);
int fts_pool_add(FtsBatch* b, const mft_fts_desc& d, int* rep);
int fts_pool_set_fields(FtsBatch* b, int rep, int pot, const float* wpl, const float* wmi);
int fts_pool_activate(FtsBatch* b, const std::vector<int>& reps);
int fts_pool_step(FtsBatch* b, int n);
hipStream_t fts_pool_stream(FtsBatch* b);
int fts_pool_observables(FtsBatch* b, int rep, double* Q, double* H, double* parts);
int fts_pool_densities(FtsBatch* b, int rep, float* rho);
int fts_pool_retire(FtsBatch* b, int rep);
int fts_pool_q_finite(FtsBatch* b, int rep, bool* ok);
int fts_pool_get_fields(FtsBatch* b, int rep, int pot, float* wpl, float* wmi);
int fts_pool_chemical_potentials(FtsBatch* b, int rep, double* mu);
size_t fts_bytes_estimate(const mft_fts_desc& d);

struct ServeReq {
    int64_t id = 0;
    int kind = 0; // 0 CL-FTS, 1 TILD
    int state = MFT_REQ_QUEUED;
    size_t bytes = 0;
    double work = 0.0;       // estimated device bytes moved by the whole request
    int passed_over = 0;     // admissions of shorter requests while this one waited
    std::string error;
    // CL-FTS
    mft_fts_desc fd;
    std::vector<std::vector<float>> wpl, wmi;
    int niter = 0, remaining = 0, rep = -1;
    std::vector<double> Q, parts, mu;
    double H[2] = {0, 0};
    std::vector<float> rho;
    std::vector<std::vector<float>> wpl_out, wmi_out; // final fields per potential
    // TILD
    mft_tild_desc td;
    std::vector<mft_tild_potential> pots;
    std::vector<float> bond_k, bond_r0, mass, mob, charge, x;
    std::vector<int> bond_style, type, nBonds, bondedTo, bondType;
    int nsteps = 0;
    mft_tild* eng = nullptr;
    hipEvent_t done = nullptr;
    std::vector<float> xout;
    std::vector<double> pe;
    double be = 0.0;
};

struct Server {
    double budget = 0.0;
    double in_use = 0.0;
    int max_reps = 0, reps_in_use = 0;
    FtsBatch* pool = nullptr;
    std::map<int64_t, std::unique_ptr<ServeReq>> reqs;
    std::deque<int64_t> queue;
    std::vector<int64_t> fts_running, tild_running;
    int64_t next_id = 1;
};

// device bytes of a TILD engine (tild_setup, configure_channels, tiles)
static size_t tild_bytes_estimate(const mft_tild_desc& d)
{
    const size_t N = d.nparticles, M = (size_t)d.nx * d.ny * d.nz;
    const size_t chan = d.nspecies + (d.has_charges ? 1 : 0);
    const size_t per_particle = 12 * 3 + 16 + 12 + 4 * 4 + 8 * (size_t)std::max(1, d.maxbonds) + 24 + 8;
    const size_t per_point = chan * (4 + 8 + 12) + (3 * chan + 1) / 2 * 8 + chan * 16 + 8;
    return N * per_particle + M * per_point + (1 << 20);
}

constexpr int kAgeLimit = 32;

static int admit(Server* s)
{
    std::vector<int> fresh;
    std::vector<int64_t> fresh_ids;
    // candidates: aged requests in arrival order, then by estimated work
    std::vector<int64_t> order(s->queue.begin(), s->queue.end());
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        const ServeReq& ra = *s->reqs[a];
        const ServeReq& rb = *s->reqs[b];
        const bool aa = ra.passed_over >= kAgeLimit, ab = rb.passed_over >= kAgeLimit;
        if (aa != ab) return aa;
        if (aa) return false; // both aged: arrival order
        return ra.work < rb.work;
    });
    std::vector<int64_t> admitted;
    for (int64_t qid : order) {
        ServeReq& r = *s->reqs[qid];
        const bool fits = s->in_use + (double)r.bytes <= s->budget || (s->in_use == 0.0);
        const bool slot = r.kind == 1 || s->reps_in_use < s->max_reps;
        if (!fits || !slot) continue;
        int rc = MFT_OK;
        if (r.kind == 0) {
            rc = fts_pool_add(s->pool, r.fd, &r.rep);
            for (int p = 0; p < r.fd.npotentials && !rc; ++p) {
                const float* a = r.wpl[p].empty() ? nullptr : r.wpl[p].data();
                const float* b = r.wmi[p].empty() ? nullptr : r.wmi[p].data();
                if (a || b) rc = fts_pool_set_fields(s->pool, r.rep, p, a, b);
            }
            if (!rc) {
                fresh.push_back(r.rep);
                fresh_ids.push_back(r.id);
                s->reps_in_use++;
            }
        } else {
            r.td.potentials = r.pots.data();
            r.td.bond_k = r.bond_k.data();
            r.td.bond_r0 = r.bond_r0.data();
            r.td.bond_style = r.bond_style.data();
            r.td.mass = r.mass.data();
            r.td.mobility = r.mob.data();
            rc = mft_tild_create(&r.eng, &r.td);
            if (!rc)
                rc = mft_tild_set_topology(r.eng, r.type.data(), r.charge.empty() ? nullptr : r.charge.data(),
                                           r.nBonds.data(), r.bondedTo.data(), r.bondType.data());
            if (!rc) rc = mft_tild_set_state(r.eng, r.x.data(), nullptr, nullptr);
            if (!rc) rc = mft_tild_step(r.eng, r.nsteps);
            if (!rc) {
                hipStream_t st = (hipStream_t)mft_tild_stream(r.eng);
                if (hipEventCreateWithFlags(&r.done, hipEventDisableTiming) != hipSuccess ||
                    hipEventRecord(r.done, st) != hipSuccess) {
                    set_error(MFT_ERR_DEVICE, "event for TILD request");
                    rc = MFT_ERR_DEVICE;
                }
            }
            if (!rc) s->tild_running.push_back(r.id);
        }
        if (rc) {
            r.state = MFT_REQ_FAILED;
            r.error = mft_last_error();
            if (r.eng) {
                mft_tild_destroy(r.eng);
                r.eng = nullptr;
            }
        } else {
            r.state = MFT_REQ_RUNNING;
            s->in_use += (double)r.bytes;
        }
        admitted.push_back(qid);
    }
    if (!admitted.empty()) {
        for (auto it = s->queue.begin(); it != s->queue.end();) {
            if (std::find(admitted.begin(), admitted.end(), *it) != admitted.end()) {
                it = s->queue.erase(it);
            } else {
                // still waiting: passed over by the requests just admitted
                s->reqs[*it]->passed_over += (int)admitted.size();
                ++it;
            }
        }
    }
    if (!fresh.empty()) {
        int rc = fts_pool_activate(s->pool, fresh);
        if (rc) return rc;
        for (int64_t id : fresh_ids) s->fts_running.push_back(id);
    }
    return MFT_OK;
}

static int run_pool(Server* s, int max_iters)
{
    if (s->fts_running.empty() || max_iters <= 0) return MFT_OK;
    int n = max_iters;
    for (int64_t id : s->fts_running) n = std::min(n, s->reqs[id]->remaining);
    int rc;
    if (n > 0) {
        if ((rc = fts_pool_step(s->pool, n))) return rc;
        for (int64_t id : s->fts_running) s->reqs[id]->remaining -= n;
    }
    MFT_HIP(hipStreamSynchronize(fts_pool_stream(s->pool)));
    // read out and retire finished solves
    std::vector<int64_t> still;
    for (int64_t id : s->fts_running) {
        ServeReq& r = *s->reqs[id];
        if (r.remaining > 0) {
            // a runaway solve ends as soon as its Q overflows, as the
            // reference quits on a NaN Hamiltonian at its next log step
            // (FTS_Box.cu:148-152), so it stops holding a pool slot
            bool ok = true;
            if ((rc = fts_pool_q_finite(s->pool, r.rep, &ok))) return rc;
            if (ok) {
                still.push_back(id);
                continue;
            }
            r.state = MFT_REQ_FAILED;
            r.error = "Found Heff = NaN, quitting! (Q not finite after " +
                      std::to_string(r.niter - r.remaining) + " iterations)";
        } else {
            const mft_fts_desc& d = r.fd;
            r.Q.assign(2 * d.nmolecules, 0.0);
            r.parts.assign(2 * (d.npotentials + d.nmolecules), 0.0);
            rc = fts_pool_observables(s->pool, r.rep, r.Q.data(), r.H, r.parts.data());
            const size_t M = (size_t)d.nx * d.ny * d.nz;
            if (!rc) {
                r.rho.resize(2 * d.nspecies * M);
                rc = fts_pool_densities(s->pool, r.rep, r.rho.data());
            }
            if (!rc) {
                r.mu.assign(2 * d.nmolecules, 0.0);
                rc = fts_pool_chemical_potentials(s->pool, r.rep, r.mu.data());
            }
            r.wpl_out.assign(d.npotentials, {});
            r.wmi_out.assign(d.npotentials, {});
            for (int p = 0; p < d.npotentials && !rc; ++p) {
                r.wpl_out[p].resize(2 * M);
                if (d.potentials[p].kind == MFT_FTS_FLORY) r.wmi_out[p].resize(2 * M);
                rc = fts_pool_get_fields(s->pool, r.rep, p, r.wpl_out[p].data(),
                                         r.wmi_out[p].empty() ? nullptr : r.wmi_out[p].data());
            }
            if (rc) {
                r.state = MFT_REQ_FAILED;
                r.error = mft_last_error();
            } else {
                r.state = MFT_REQ_DONE;
            }
        }
        if ((rc = fts_pool_retire(s->pool, r.rep))) return rc;
        r.rep = -1;
        s->reps_in_use--;
        s->in_use -= (double)r.bytes;
    }
    s->fts_running.swap(still);
    return MFT_OK;
}

static int poll_tild(Server* s)
{
    std::vector<int64_t> still;
    for (int64_t id : s->tild_running) {
        ServeReq& r = *s->reqs[id];
        hipError_t q = hipEventQuery(r.done);
        if (q == hipErrorNotReady) {
            still.push_back(id);
            continue;
        }
        int rc = q == hipSuccess ? MFT_OK : MFT_ERR_DEVICE;
        if (!rc) {
            r.xout.resize(3 * (size_t)r.td.nparticles);
            rc = mft_tild_get_state(r.eng, r.xout.data(), nullptr, nullptr);
        }
        if (!rc) {
            r.pe.assign(std::max(1, r.td.npotentials), 0.0);
            rc = mft_tild_energies(r.eng, r.pe.data(), &r.be);
        }
        r.state = rc ? MFT_REQ_FAILED : MFT_REQ_DONE;
        if (rc) r.error = q == hipSuccess ? mft_last_error() : hipGetErrorString(q);
        (void)hipEventDestroy(r.done);
        r.done = nullptr;
        mft_tild_destroy(r.eng);
        r.eng = nullptr;
        s->in_use -= (double)r.bytes;
    }
    s->tild_running.swap(still);
    return MFT_OK;
}

} // namespace mft

using namespace mft;

struct mft_server : Server {};

extern "C" {

int mft_server_create(mft_server** out, double budget, int max_reps)
{
    MFT_REQUIRE(out && budget > 0.0 && max_reps >= 1, MFT_ERR_INVALID, "bad argument");
    mft_server* s = new mft_server();
    s->budget = budget;
    s->max_reps = max_reps;
    int rc = fts_pool_create(&s->pool, max_reps);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return MFT_OK;
}

int mft_server_destroy(mft_server* s)
{
    if (!s) return MFT_OK;
    for (auto& kv : s->reqs) {
        ServeReq& r = *kv.second;
        if (r.eng) {
            (void)mft_tild_synchronize(r.eng);
            mft_tild_destroy(r.eng);
        }
        if (r.done) (void)hipEventDestroy(r.done);
    }
    fts_pool_destroy(s->pool);
    delete s;
    return MFT_OK;
}

int mft_server_submit_fts(mft_server* s, const mft_fts_desc* desc, const float* const* wpl, const float* const* wmi,
                          int niter, int64_t* id)
{
    MFT_REQUIRE(s && desc && id && niter >= 0, MFT_ERR_INVALID, "bad argument");
    MFT_REQUIRE(desc->npotentials >= 0 && desc->npotentials <= 8, MFT_ERR_INVALID, "too many potentials");
    auto r = std::make_unique<ServeReq>();
    r->kind = 0;
    r->fd = *desc;
    r->niter = r->remaining = niter;
    const size_t M = (size_t)desc->nx * desc->ny * desc->nz;
    r->wpl.resize(desc->npotentials);
    r->wmi.resize(desc->npotentials);
    for (int p = 0; p < desc->npotentials; ++p) {
        if (wpl && wpl[p]) r->wpl[p].assign(wpl[p], wpl[p] + 2 * M);
        if (wmi && wmi[p]) r->wmi[p].assign(wmi[p], wmi[p] + 2 * M);
    }
    r->bytes = fts_bytes_estimate(*desc);
    {
        // one CL iteration moves about M * 8 * (8 Ns + 12) bytes per molecule
        double w = 0.0;
        for (int m = 0; m < desc->nmolecules; ++m) {
            int ns = 0;
            for (int b = 0; b < desc->molecules[m].nblocks; ++b) ns += desc->molecules[m].N[b];
            w += (double)M * 8.0 * (8.0 * ns + 12.0);
        }
        r->work = w * (double)niter;
    }
    r->id = s->next_id++;
    *id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    return MFT_OK;
}

int mft_server_submit_tild(mft_server* s, const mft_tild_desc* desc, const int* type, const float* charge,
                           const int* nBonds, const int* bondedTo, const int* bondType, const float* x, int nsteps,
                           int64_t* id)
{
    MFT_REQUIRE(s && desc && type && x && id && nsteps >= 0 && desc->nparticles > 0, MFT_ERR_INVALID,
                "bad argument");
    auto r = std::make_unique<ServeReq>();
    r->kind = 1;
    r->td = *desc;
    const size_t N = desc->nparticles, mb = std::max(1, desc->maxbonds);
    r->pots.assign(desc->potentials, desc->potentials + std::max(0, desc->npotentials));
    r->bond_k.assign(desc->bond_k, desc->bond_k + std::max(0, desc->nbondtypes));
    r->bond_r0.assign(desc->bond_r0, desc->bond_r0 + std::max(0, desc->nbondtypes));
    r->bond_style.assign(desc->bond_style, desc->bond_style + std::max(0, desc->nbondtypes));
    r->mass.assign(desc->nspecies, 1.0f);
    r->mob.assign(desc->nspecies, 1.0f);
    for (int k = 0; k < desc->nspecies; ++k) {
        if (desc->mass) r->mass[k] = desc->mass[k];
        if (desc->mobility) r->mob[k] = desc->mobility[k];
    }
    r->type.assign(type, type + N);
    if (charge) r->charge.assign(charge, charge + N);
    r->nBonds.assign(N, 0);
    r->bondedTo.assign(N * mb, 0);
    r->bondType.assign(N * mb, 0);
    if (nBonds) {
        r->nBonds.assign(nBonds, nBonds + N);
        r->bondedTo.assign(bondedTo, bondedTo + N * mb);
        r->bondType.assign(bondType, bondType + N * mb);
    }
    r->x.assign(x, x + 3 * N);
    r->nsteps = nsteps;
    r->bytes = tild_bytes_estimate(*desc);
    {
        // one step moves about N * 112 + M * channels * 80 bytes
        const double M = (double)desc->nx * desc->ny * desc->nz;
        const double chan = desc->nspecies + (desc->has_charges ? 1 : 0);
        r->work = ((double)N * 112.0 + M * chan * 80.0) * (double)nsteps;
    }
    r->id = s->next_id++;
    *id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    return MFT_OK;
}

int mft_server_pump(mft_server* s, int max_iters, int* nrunning, int* nqueued)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null server");
    int rc;
    if ((rc = admit(s))) return rc;
    if ((rc = poll_tild(s))) return rc;
    if ((rc = run_pool(s, max_iters))) return rc;
    if ((rc = poll_tild(s))) return rc;
    if (nrunning) *nrunning = (int)(s->fts_running.size() + s->tild_running.size());
    if (nqueued) *nqueued = (int)s->queue.size();
    return MFT_OK;
}

int mft_server_status(mft_server* s, int64_t id, int* state)
{
    MFT_REQUIRE(s && state, MFT_ERR_INVALID, "bad argument");
    auto it = s->reqs.find(id);
    MFT_REQUIRE(it != s->reqs.end(), MFT_ERR_INVALID, "unknown request %lld", (long long)id);
    *state = it->second->state;
    return MFT_OK;
}

const char* mft_server_error(mft_server* s, int64_t id)
{
    if (!s) return "null server";
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return "unknown request";
    return it->second->error.c_str();
}

int mft_server_fts_result(mft_server* s, int64_t id, double* Q, double* H, double* parts, float* rho)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null server");
    auto it = s->reqs.find(id);
    MFT_REQUIRE(it != s->reqs.end() && it->second->kind == 0, MFT_ERR_INVALID, "no CL-FTS request %lld",
                (long long)id);
    ServeReq& r = *it->second;
    MFT_REQUIRE(r.state != MFT_REQ_FAILED, MFT_ERR_DEVICE, "request failed: %s", r.error.c_str());
    MFT_REQUIRE(r.state == MFT_REQ_DONE, MFT_ERR_INVALID, "request %lld not finished", (long long)id);
    if (Q) std::copy(r.Q.begin(), r.Q.end(), Q);
    if (H) {
        H[0] = r.H[0];
        H[1] = r.H[1];
    }
    if (parts) std::copy(r.parts.begin(), r.parts.end(), parts);
    if (rho) std::copy(r.rho.begin(), r.rho.end(), rho);
    return MFT_OK;
}

int mft_server_fts_final_state(mft_server* s, int64_t id, int pot, float* wpl, float* wmi, double* mu)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null server");
    auto it = s->reqs.find(id);
    MFT_REQUIRE(it != s->reqs.end() && it->second->kind == 0, MFT_ERR_INVALID, "no CL-FTS request %lld",
                (long long)id);
    ServeReq& r = *it->second;
    MFT_REQUIRE(r.state != MFT_REQ_FAILED, MFT_ERR_DEVICE, "request failed: %s", r.error.c_str());
    MFT_REQUIRE(r.state == MFT_REQ_DONE, MFT_ERR_INVALID, "request %lld not finished", (long long)id);
    MFT_REQUIRE(pot >= 0 && pot < r.fd.npotentials, MFT_ERR_INVALID, "potential out of range");
    if (wpl) std::copy(r.wpl_out[pot].begin(), r.wpl_out[pot].end(), wpl);
    if (wmi) {
        MFT_REQUIRE(!r.wmi_out[pot].empty(), MFT_ERR_INVALID, "potential %d has no w- field", pot);
        std::copy(r.wmi_out[pot].begin(), r.wmi_out[pot].end(), wmi);
    }
    if (mu) std::copy(r.mu.begin(), r.mu.end(), mu);
    return MFT_OK;
}

int mft_server_tild_result(mft_server* s, int64_t id, float* x, double* pot_energy, double* bond_energy)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null server");
    auto it = s->reqs.find(id);
    MFT_REQUIRE(it != s->reqs.end() && it->second->kind == 1, MFT_ERR_INVALID, "no TILD request %lld",
                (long long)id);
    ServeReq& r = *it->second;
    MFT_REQUIRE(r.state != MFT_REQ_FAILED, MFT_ERR_DEVICE, "request failed: %s", r.error.c_str());
    MFT_REQUIRE(r.state == MFT_REQ_DONE, MFT_ERR_INVALID, "request %lld not finished", (long long)id);
    if (x) std::copy(r.xout.begin(), r.xout.end(), x);
    if (pot_energy) std::copy(r.pe.begin(), r.pe.begin() + r.td.npotentials, pot_energy);
    if (bond_energy) *bond_energy = r.be;
    return MFT_OK;
}

int mft_server_release(mft_server* s, int64_t id)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null server");
    auto it = s->reqs.find(id);
    MFT_REQUIRE(it != s->reqs.end(), MFT_ERR_INVALID, "unknown request %lld", (long long)id);
    MFT_REQUIRE(it->second->state == MFT_REQ_DONE || it->second->state == MFT_REQ_FAILED, MFT_ERR_INVALID,
                "request %lld still queued or running", (long long)id);
    s->reqs.erase(it);
    return MFT_OK;
}

} // extern "C"
