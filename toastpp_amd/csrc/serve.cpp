// Serving loop: continuous batching of CalcFields requests on one GPU.
//
// A request is a prepared solver (mesh + assembled matrix, resident on the
// device) and a block of right-hand sides in device memory.  The server owns
// `slots` worker threads, each with its own HIP stream and solve workspace
// (Solver::Exec).  Requests are admitted in arrival order as soon as a slot
// is free and the device memory the server's workspaces may hold fits the
// HBM budget; admitted requests run concurrently, their kernels
// interleaving on the GPU.  All reductions are fixed-order, so a request's
// result does not depend on what else is running.
//
// Accounting: a slot keeps the workspaces of its earlier requests (reused
// when the next request has the same mesh size and tile widths).  Each slot
// counts held (the bytes its workspace holds between requests) plus
// reserved (the bound Solver::work_bytes gives for its running request,
// FP64 fallback included), and a request is admitted on a slot only when
// the sum over all slots stays within the budget.  An idle slot whose held
// memory keeps the queue's head out frees its workspace.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "handles.hpp"

namespace tmi {

void set_last_error(const std::string &msg);

struct ServeRequest {
    Solver *solver = nullptr;
    int nq = 0, ld = 0;
    const void *d_b = nullptr;
    void *d_x = nullptr;
    size_t bytes = 0;
    double t_queued = 0, t_started = 0, t_finished = 0;
    int status = TMI_OK;
    bool done = false;
    std::string error;
    std::vector<tmi_colstat> stats;
};

class Server {
   public:
    Server(int slots, double budget)
        : budget_(budget), held_(slots, 0.0), resv_(slots, 0.0), t0_(std::chrono::steady_clock::now())
    {
        for (int i = 0; i < slots; i++) execs_.emplace_back(new Solver::Exec(true));
        for (int i = 0; i < slots; i++) workers_.emplace_back(&Server::work, this, i);
    }
    ~Server()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_work_.notify_all();
        for (auto &t : workers_) t.join();
    }
    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }
    long long submit(Solver *s, int nq, const void *d_b, void *d_x, int ld)
    {
        auto r = std::make_unique<ServeRequest>();
        r->solver = s;
        r->nq = nq;
        r->ld = ld;
        r->d_b = d_b;
        r->d_x = d_x;
        r->bytes = s->work_bytes(nq);
        TMI_CHECK((double)r->bytes <= budget_, TMI_ERR_OOM,
                  "request needs %.1f MiB of workspace, more than the server's HBM budget", r->bytes / 1048576.0);
        r->t_queued = now();
        std::lock_guard<std::mutex> lk(mu_);
        long long id = next_id_++;
        reqs_[id] = std::move(r);
        queue_.push_back(id);
        cv_work_.notify_all();
        return id;
    }
    // wait for request id (timeout_s < 0: forever); returns false on timeout
    bool wait(long long id, double timeout_s, ServeRequest &out)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        TMI_CHECK(it != reqs_.end(), TMI_ERR_INVALID_ARG, "unknown request id %lld", id);
        auto pred = [&] { return it->second->done; };
        if (timeout_s < 0) cv_done_.wait(lk, pred);
        else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return false;
        out = *it->second;
        reqs_.erase(it);
        return true;
    }
    void counts(int *running, int *queued, double *bytes)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (running) *running = running_;
        if (queued) *queued = (int)queue_.size();
        if (bytes) *bytes = accounted();
    }

   private:
    // device bytes counted for all slots (caller holds mu_)
    double accounted() const
    {
        double b = 0;
        for (size_t j = 0; j < held_.size(); j++) b += held_[j] + resv_[j];
        return b;
    }
    // whether a request of `need` bytes fits on idle slot `slot` (caller holds mu_)
    bool fits(int slot, double need) const { return accounted() + need <= budget_; }

    void work(int slot)
    {
        Solver::Exec &e = *execs_[slot];
        for (;;) {
            ServeRequest *r = nullptr;
            bool trim = false;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_work_.wait(lk, [&] {
                    if (stop_) return true;
                    if (queue_.empty()) return false;
                    // admit, or give back this idle slot's workspace if it keeps the head out
                    return fits(slot, reqs_[queue_.front()]->bytes) || held_[slot] > 0;
                });
                if (stop_ && queue_.empty()) return;
                if (queue_.empty()) continue;
                long long id = queue_.front();
                if (!fits(slot, reqs_[id]->bytes)) {
                    trim = true;
                } else {
                    queue_.pop_front();
                    r = reqs_[id].get();
                    resv_[slot] = (double)r->bytes;
                    running_++;
                    r->t_started = now();
                }
            }
            if (trim) {
                try {
                    e.release_work();
                } catch (...) {
                }
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    held_[slot] = (double)e.bytes();
                }
                cv_work_.notify_all();
                continue;
            }
            r->stats.assign(r->nq, tmi_colstat{});
            try {
                int nc = r->solver->solve_on(e, r->nq, r->d_b, r->d_x, r->ld, r->stats.data());
                r->status = nc ? TMI_ERR_NOT_CONVERGED : TMI_OK;
            } catch (const Error &ex) {
                r->status = ex.code;
                r->error = ex.what();
            } catch (const std::exception &ex) {
                r->status = TMI_ERR_DEVICE;
                r->error = ex.what();
            }
            const double held = (double)e.bytes();
            {
                std::lock_guard<std::mutex> lk(mu_);
                r->t_finished = now();
                r->done = true;
                held_[slot] = held;
                resv_[slot] = 0;
                running_--;
            }
            cv_done_.notify_all();
            cv_work_.notify_all();
        }
    }

    double budget_;
    std::vector<double> held_, resv_;   // per slot: workspace kept between requests, bound of the running request
    int running_ = 0;
    bool stop_ = false;
    long long next_id_ = 1;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<long long> queue_;
    std::unordered_map<long long, std::unique_ptr<ServeRequest>> reqs_;
    std::vector<std::unique_ptr<Solver::Exec>> execs_;
    std::vector<std::thread> workers_;
};

}  // namespace tmi

struct tmi_server_s {
    std::unique_ptr<tmi::Server> srv;
};

using namespace tmi;

#define API_BEGIN try {
#define API_END                                   \
    return TMI_OK;                                \
    }                                             \
    catch (const tmi::Error &e)                   \
    {                                             \
        set_last_error(e.what());                 \
        return e.code;                            \
    }                                             \
    catch (const std::exception &e)               \
    {                                             \
        set_last_error(e.what());                 \
        return TMI_ERR_INVALID_ARG;               \
    }

extern "C" {

int tmi_server_create(int slots, double hbm_budget_bytes, tmi_server *out)
{
    API_BEGIN
    TMI_CHECK(out && slots >= 1 && slots <= 16 && hbm_budget_bytes > 0, TMI_ERR_INVALID_ARG,
              "server needs 1..16 slots and a positive HBM budget");
    Context::get();
    tmi_server s = new tmi_server_s;
    s->srv.reset(new Server(slots, hbm_budget_bytes));
    *out = s;
    API_END
}

int tmi_server_free(tmi_server s)
{
    API_BEGIN
    delete s;
    API_END
}

int tmi_server_submit(tmi_server s, tmi_solver solver, int nq, const void *d_rhs, void *d_phi, int ld, long long *id)
{
    API_BEGIN
    TMI_CHECK(s && solver && d_rhs && d_phi && id && nq > 0 && ld >= nq, TMI_ERR_INVALID_ARG, "invalid request");
    Solver &sv = solver->s;
    TMI_CHECK(sv.assembled, TMI_ERR_STATE, "request on a solver that has not been Reset");
    *id = s->srv->submit(&sv, nq, d_rhs, d_phi, ld);
    API_END
}

int tmi_server_wait(tmi_server s, long long id, double timeout_s, tmi_colstat *stats, double *times, int *status)
{
    API_BEGIN
    TMI_CHECK(s, TMI_ERR_INVALID_ARG, "invalid server handle");
    ServeRequest r;
    if (!s->srv->wait(id, timeout_s, r)) return TMI_ERR_TIMEOUT;   // still queued or running
    if (stats) std::copy(r.stats.begin(), r.stats.end(), stats);
    if (times) { times[0] = r.t_queued; times[1] = r.t_started; times[2] = r.t_finished; }
    if (status) *status = r.status;
    if (r.status != TMI_OK && r.status != TMI_ERR_NOT_CONVERGED) set_last_error(r.error);
    API_END
}

int tmi_server_counts(tmi_server s, int *running, int *queued, double *hbm_bytes)
{
    API_BEGIN
    TMI_CHECK(s, TMI_ERR_INVALID_ARG, "invalid server handle");
    s->srv->counts(running, queued, hbm_bytes);
    API_END
}

double tmi_server_clock(tmi_server s) { return s ? s->srv->now() : 0.0; }

}  // extern "C"
