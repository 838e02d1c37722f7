// DynamicInputRecord: an InputRecord built in memory (the reference's
// src/core/dynamicinputrecord.h:59-135: setField per type, giveField from
// the matching map with an int -> double fallback, MissingKeywordInputException
// for an absent keyword, dynamicinputrecord.C:166-282).  The engineering
// model fills one with the linear-solver keywords of the run and hands it
// to the solver's initializeFrom, as NRSolver forwards its own record
// (nrsolver.C:130-132).
#pragma once

#include <map>
#include <set>

#include "inputrecord.h"

namespace oofem {

class DynamicInputRecord : public InputRecord {
    std::string recordKeyword;
    std::set<std::string> emptyRecord;
    std::map<std::string, int> intRecord;
    std::map<std::string, double> doubleRecord;
    std::map<std::string, bool> boolRecord;
    std::map<std::string, std::string> stringRecord;
    std::map<std::string, FloatArray> floatArrayRecord;
    std::map<std::string, IntArray> intArrayRecord;

    template <typename M>
    static auto &find(M &m, InputFieldType id, const InputRecord &ir)
    {
        auto it = m.find(id);
        if (it == m.end()) throw MissingKeywordInputException(ir, id, -1);
        return it->second;
    }

public:
    explicit DynamicInputRecord(std::string keyword = "") : recordKeyword(std::move(keyword)) {}

    std :: string giveRecordAsString() const override;
    std :: string giveLocation() const override { return "record \"" + recordKeyword + "\""; }

    void giveField(int &answer, InputFieldType id) override { answer = find(intRecord, id, *this); }
    void giveField(double &answer, InputFieldType id) override
    {
        auto it = doubleRecord.find(id);
        if (it != doubleRecord.end()) {
            answer = it->second;
            return;
        }
        answer = find(intRecord, id, *this);
    }
    void giveField(bool &answer, InputFieldType id) override { answer = find(boolRecord, id, *this); }
    void giveField(std :: string &answer, InputFieldType id) override { answer = find(stringRecord, id, *this); }
    void giveField(FloatArray &answer, InputFieldType id) override { answer = find(floatArrayRecord, id, *this); }
    void giveField(IntArray &answer, InputFieldType id) override { answer = find(intArrayRecord, id, *this); }

    bool hasField(InputFieldType id) override
    {
        return emptyRecord.count(id) || intRecord.count(id) || doubleRecord.count(id) || boolRecord.count(id) ||
               stringRecord.count(id) || floatArrayRecord.count(id) || intArrayRecord.count(id);
    }

    void setField(int item, InputFieldType id) { intRecord[id] = item; }
    void setField(double item, InputFieldType id) { doubleRecord[id] = item; }
    void setField(bool item, InputFieldType id) { boolRecord[id] = item; }
    void setField(std :: string item, InputFieldType id) { stringRecord[id] = std::move(item); }
    void setField(FloatArray item, InputFieldType id) { floatArrayRecord[id] = std::move(item); }
    void setField(IntArray item, InputFieldType id) { intArrayRecord[id] = std::move(item); }
    void setField(InputFieldType id) { emptyRecord.insert(id); }
    void unsetField(InputFieldType id)
    {
        emptyRecord.erase(id);
        intRecord.erase(id);
        doubleRecord.erase(id);
        boolRecord.erase(id);
        stringRecord.erase(id);
        floatArrayRecord.erase(id);
        intArrayRecord.erase(id);
    }
};

inline std :: string DynamicInputRecord::giveRecordAsString() const
{
    std::string s = recordKeyword;
    for (auto &k : emptyRecord) s += " " + k;
    for (auto &kv : intRecord) s += " " + kv.first + " " + std::to_string(kv.second);
    for (auto &kv : doubleRecord) {
        char buf[64];
        snprintf(buf, sizeof buf, "%.17g", kv.second);
        s += " " + kv.first + " " + buf;
    }
    for (auto &kv : boolRecord) s += " " + kv.first + " " + (kv.second ? "1" : "0");
    for (auto &kv : stringRecord) s += " " + kv.first + " \"" + kv.second + "\"";
    return s;
}

}  // namespace oofem
