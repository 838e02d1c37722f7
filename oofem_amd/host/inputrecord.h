// InputRecord: the reference's keyword/value record interface
// (src/core/inputrecord.h:59-75 InputFieldType and the IR_GIVE_* macros,
// :88-231 the giveField / giveOptionalField / hasField members, :262-282 the
// input exceptions), restated with the same names and signatures for the
// field types the plug-ins read, so that HipPCGSolver::initializeFrom reads
// its keywords exactly as IMLSolver::initializeFrom does (imlsolver.C:67-97).
#pragma once

#include <stdexcept>
#include <string>

#include "oofem_types.h"

namespace oofem {

typedef const char *InputFieldType;

#define IR_GIVE_FIELD(__ir, __value, __id) (__ir).giveField(__value, __id);
#define IR_GIVE_OPTIONAL_FIELD(__ir, __value, __id) (__ir).giveOptionalField(__value, __id);

class InputRecord {
public:
    InputRecord() {}
    virtual ~InputRecord() = default;

    virtual std :: string giveRecordAsString() const = 0;
    virtual std :: string giveLocation() const = 0;

    virtual void giveField(int &answer, InputFieldType id) = 0;
    virtual void giveField(double &answer, InputFieldType id) = 0;
    virtual void giveField(bool &answer, InputFieldType id) = 0;
    virtual void giveField(std :: string &answer, InputFieldType id) = 0;
    virtual void giveField(FloatArray &answer, InputFieldType id) = 0;
    virtual void giveField(IntArray &answer, InputFieldType id) = 0;

    template<typename T>
    void giveOptionalField(T& answer, InputFieldType id){ if(hasField(id)) giveField(answer,id); }

    virtual bool hasField(InputFieldType id) = 0;
};

// input errors carry the record and the keyword (inputrecord.h:249-282);
// they derive from RuntimeException here so that callers catching the
// plug-ins' OOFEM_ERROR see them too
class InputException : public RuntimeException {
public:
    std::string record;
    std::string keyword;
    int number;
    InputException(const InputRecord &ir, std::string kw, int num, const std::string &what) :
        RuntimeException(what), record(ir.giveRecordAsString()), keyword(std::move(kw)), number(num)
    {}
};

class MissingKeywordInputException : public InputException {
public:
    MissingKeywordInputException(const InputRecord &ir, std::string keyword, int number) :
        InputException(ir, keyword, number, "Missing keyword \"" + keyword + "\" in " + ir.giveLocation())
    {}
};

class BadFormatInputException : public InputException {
public:
    BadFormatInputException(const InputRecord &ir, std::string keyword, int number) :
        InputException(ir, keyword, number, "Bad format for keyword \"" + keyword + "\" in " + ir.giveLocation())
    {}
};

class ValueInputException : public InputException {
public:
    ValueInputException(const InputRecord &ir, std::string keyword, const std::string &reason) :
        InputException(ir, keyword, -1,
                       "Bad value for keyword \"" + keyword + "\" in " + ir.giveLocation() + ": " + reason)
    {}
};

}  // namespace oofem
