// Reference-compatible driver: `als_main <config.ini>` reads the same INI
// keys as the reference's main (main.cpp:21-47), loads the reference .bin
// files (main.cpp:68-75), runs als_runs trainings (main.cpp:82-114), prints
// the final test RMSE and writes the CSV event log of logger.cpp:71
// (columns log_name,id,run_iter,als_iter,prev_ts,curr_ts,elapsed,
// event_elapsed,text; events ALS_UPDATE_U/V, ALS_ITER, MF_MODEL_TRAINING).
// Extension keys: [als] gram_prec (0 FP32, 1 BF16, 2 FP16), refine_steps;
// [general] model_path (libmf dump after a single run).
#include <cctype>
#include <chrono>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

#include "als_model.hpp"

namespace {

std::map<std::string, std::string> parse_ini(const std::string &path, bool &ok) {
    std::map<std::string, std::string> kv;
    std::ifstream in(path);
    ok = (bool)in;
    std::string line, section;
    auto trim = [](std::string s) {
        size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    };
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (line[0] == '[') {
            section = trim(line.substr(1, line.find(']') - 1));
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        size_t c = val.find_first_of(";#");
        if (c != std::string::npos) val = trim(val.substr(0, c));
        kv[section + "." + key] = val;
    }
    return kv;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch()).count();
}

struct CsvLog {
    std::string name;
    std::stringstream ss;
    double prev = 0;
    size_t id = 0, run_iter = 0, als_iter = 0;
    void init() {
        std::time_t t = std::time(nullptr);
        char buf[42];
        std::strftime(buf, sizeof(buf), "%Y%m%d_%H%M%S", std::gmtime(&t));
        name = std::string(buf) + "_debug=false.csv";
        prev = now_s();
        ss << std::fixed << "log_name,id,run_iter,als_iter,prev_ts,curr_ts,elapsed,event_elapsed,text\n";
        log("init logger", false);
    }
    void log(const std::string &text, bool cout, double event_elapsed = -1) {
        double cur = now_s(), el = cur - prev;
        prev = cur;
        ss << name << "," << id++ << "," << run_iter << "," << als_iter << "," << prev << "," << cur << "," << el << ",";
        if (event_elapsed >= 0) ss << event_elapsed;
        ss << "," << text << "\n";
        if (cout) std::cout << "run iter: " << run_iter << " als iter: " << als_iter << " elapsed: " << el << " " << text << std::endl;
    }
    void save(const std::string &folder) {
        if (folder.empty()) return;
        std::ofstream out(folder + "/" + name);
        out << ss.str();
    }
};

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <config.ini>\n";
        return 1;
    }
    bool ok = false;
    auto kv = parse_ini(argv[1], ok);
    if (!ok) {
        std::cout << "Can't load " << argv[1] << "\n";
        return 1;
    }
    auto geti = [&](const std::string &k, long d) { auto it = kv.find(k); return it == kv.end() ? d : std::stol(it->second); };
    auto getf = [&](const std::string &k, double d) { auto it = kv.find(k); return it == kv.end() ? d : std::stod(it->second); };
    auto gets = [&](const std::string &k, const std::string &d) { auto it = kv.find(k); return it == kv.end() ? d : it->second; };

    const int m = (int)geti("data.m", 0), n = (int)geti("data.n", 0), f = (int)geti("als.f", 0);
    const long nnz_train = geti("data.nnz_train", 0), nnz_test = geti("data.nnz_test", 0);
    const float lambda = (float)getf("als.lambda", 0);
    const int iters = (int)geti("als.als_iters", 0), runs = (int)geti("general.als_runs", 0);
    const std::string data = gets("data.data_folder", ""), log_folder = gets("general.log_folder", "");
    const int vtvs = (int)geti("als.als_calculate_vtvs_type", 0), solve = (int)geti("als.als_solve_type", 0);
    const int mb = (int)geti("als.m_batches", 1), nb = (int)geti("als.n_batches", 1);

    als_config cfg;
    if (als_config_from_reference(f, lambda, iters, vtvs, solve, 32, mb, nb, &cfg) != ALS_OK) {
        std::cerr << "invalid configuration\n";
        return 1;
    }
    cfg.gram_prec = (int)geti("als.gram_prec", cfg.gram_prec);
    cfg.refine_steps = (int)geti("als.refine_steps", 0);

    CsvLog lg;
    lg.init();
    try {
        host_sparse_matrix tr(m, n, nnz_train), te(m, n, nnz_test);
        tr.load_csc(data + "/R_train_csc.data.bin", data + "/R_train_csc.indices.bin", data + "/R_train_csc.indptr.bin");
        tr.load_csr(data + "/R_train_csr.data.bin", data + "/R_train_csr.indptr.bin", data + "/R_train_csr.indices.bin");
        te.load_coo(data + "/R_test_coo.data.bin", data + "/R_test_coo.row.bin", data + "/R_test_coo.col.bin");
        lg.log("host train and test ratings loaded", true);
        for (int r = 0; r < runs; r++) {
            lg.run_iter = r + 1;
            double t_train = now_s();
            als_model model(tr, te, cfg);
            lg.log("als model constructor done", true);
            if (als_init_factors(model.engine, ALS_INIT_REFERENCE) != ALS_OK) throw std::runtime_error(als_last_error(model.engine));
            for (int it = 0; it < iters; it++) {
                lg.als_iter = it + 1;
                double t_it = now_s(), t0 = now_s();
                model.update_U();
                als_synchronize(model.engine);
                lg.log("event ALS_UPDATE_U finished", true, now_s() - t0);
                t0 = now_s();
                model.update_V();
                als_synchronize(model.engine);
                lg.log("event ALS_UPDATE_V finished", true, now_s() - t0);
                lg.log("event ALS_ITER finished", true, now_s() - t_it);
                // the reference's CALC_ITER_RMSE report (als_model.cu:2269-2274)
                if (it < iters - 1) lg.log("RMSE test: " + std::to_string(model.rsme_test()), true);
            }
            lg.als_iter = 0;
            if (als_get_factors(model.engine, ALS_SIDE_USERS, model.h_UT, m, f) != ALS_OK ||
                als_get_factors(model.engine, ALS_SIDE_ITEMS, model.h_VT, n, f) != ALS_OK)
                throw std::runtime_error(als_last_error(model.engine));
            lg.log("event MF_MODEL_TRAINING finished", true, now_s() - t_train);
            lg.log("final RMSE test: " + std::to_string(model.rsme_test()), true);
            std::string mp = gets("general.model_path", "");
            if (runs == 1 && !mp.empty()) {
                model.save_libmf(mp.c_str());
                lg.log("saved model in cumf_sgd libmf format", true);
            }
        }
    } catch (const std::exception &ex) {
        std::cerr << "error: " << ex.what() << "\n";
        lg.save(log_folder);
        return 1;
    }
    lg.run_iter = 0;
    lg.log("all als runs done", true);
    lg.save(log_folder);
    return 0;
}
